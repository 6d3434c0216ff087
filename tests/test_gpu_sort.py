"""GpuSort on the GPU (src/bvh/sort.rs:13-111) == the oracle's sort by (code, slot).

Codes of all 32 bits, the five sets of src/bvh/sort.rs:156-175 (golden, from
tests/golden/golden.json), sizes across tile boundaries, and a replay after
new codes are written into the same array (the recorded sort is reused, as
GpuBvh::update reuses it). Bit-exact: order and sorted codes.
"""
import json
import pathlib

import numpy as np
import pytest

import phiray_amd as P

pytestmark = pytest.mark.gpu
GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "golden.json").read_text())


def check(gpu, orc, codes):
    arr = P.GpuArray.from_(gpu, np.ascontiguousarray(codes, dtype=np.uint32))
    s = P.GpuSort(gpu, arr)
    s.run()
    order, srt = s.read()
    want_o, want_s = orc.sort_codes(np.ascontiguousarray(codes, dtype=np.uint32))
    assert (order == want_o).all() and (srt == want_s).all()
    keys = s.keys()
    assert (keys[:, 0] == want_o).all() and (keys[:, 1] == want_s).all()
    return arr, s


@pytest.mark.parametrize("n", [1, 2, 2047, 2048, 2049, 100_003])
def test_random_full_width_codes(gpu, orc, n):
    print(f"[sort] n={n}", flush=True)
    rng = np.random.default_rng(n)
    check(gpu, orc, rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32))


def test_reference_sort_sets(gpu, orc):
    print("[sort] reference sets", flush=True)
    for case in GOLDEN["sort_sets"]:
        check(gpu, orc, np.array(case["codes"], dtype=np.uint32))


def test_replay_after_new_codes(gpu, orc):
    print("[sort] replay", flush=True)
    rng = np.random.default_rng(7)
    a = rng.integers(0, 1000, 50_000).astype(np.uint32)  # many duplicates: stability matters
    arr, s = check(gpu, orc, a)
    b = rng.integers(0, 2 ** 32, 50_000, dtype=np.uint64).astype(np.uint32)
    arr.write(b)
    s.run()
    order, srt = s.read()
    want_o, want_s = orc.sort_codes(b)
    assert (order == want_o).all() and (srt == want_s).all()
    assert s.steps() == 6 and len(s) == 50_000
