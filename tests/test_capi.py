"""The C-ABI library loads and exports every entry point include/phiray_hip.h
declares (CPU only: no compute call is made without a device)."""
import ctypes
import pathlib
import re

import phiray_amd as P

HEADER = pathlib.Path(__file__).resolve().parent.parent / "include" / "phiray_hip.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(phr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ["phr_gpu_open", "phr_bvh_build", "phr_bvh_update", "phr_bvh_download", "phr_tracer_new",
                 "phr_tracer_sample", "phr_tracer_raycast", "phr_array_read_range"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(P.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_the_header():
    assert set(declared()) == set(P.exported_symbols())


def test_abi_version_and_error_slot():
    L = P.lib()
    assert L.phr_abi_version() == 1
    assert isinstance(L.phr_last_error(), bytes)
