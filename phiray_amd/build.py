"""Build the gfx950 HIP library in-tree: phiray_amd/libphiray_hip.so.

One hipcc invocation over every .hip file in csrc/. The flags matter for
correctness, not only speed:

* ``-ffp-contract=off``: no a*b+c is fused, so every expression rounds as the
  CPU twin's does (the watertight triangle test depends on it);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32
  instructions are generated;
* correctly rounded f32 division and sqrt are hipcc's defaults and are kept.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libphiray_hip.so"
ARCH = os.environ.get("PHIRAY_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and pathlib.Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-function",
        f"-I{CSRC}",
        f"-I{HERE.parent / 'include'}",
    ]


def build(verbose: bool = False) -> pathlib.Path:
    sources = sorted(str(p) for p in CSRC.glob("*.hip"))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *flags(), *sources, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libphiray_hip.so")
    noisy = [ln for ln in res.stderr.splitlines() if "warning" in ln and "packed-fp32-ops" not in ln]
    if verbose and noisy:
        sys.stderr.write("\n".join(noisy) + "\n")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
