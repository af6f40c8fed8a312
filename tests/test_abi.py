"""The C-ABI library loads, exports every declared entry point, and agrees
with the Python struct mirrors on sizes (no device calls)."""
import ctypes as C
import os
import re

import pytest

from jurassic_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(\w+)\s*\(", src, re.M)
    return sorted(set(names))


def test_exports_all_declared_symbols():
    from jurassic_amd import gpu
    lib = C.CDLL(gpu.LIBPATH)
    names = declared_functions()
    assert "formod_gpu" in names and "jur_ctx_formod_device" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(gpu.EXPORTS) <= set(names)


def test_struct_sizes_match_library():
    from jurassic_amd import gpu
    lib = C.CDLL(gpu.LIBPATH)
    lib.jur_gpu_abi_check.argtypes = [C.c_size_t] * 3
    assert lib.jur_gpu_abi_check(C.sizeof(abi.ctl_t), C.sizeof(abi.atm_t), C.sizeof(abi.obs_t)) == 0
    assert lib.jur_gpu_abi_check(C.sizeof(abi.ctl_t) + 8, C.sizeof(abi.atm_t), C.sizeof(abi.obs_t)) == 1


def test_struct_layout_of_reference_fields():
    # field order of src/jurassic.h: offsets follow from the capacities
    assert abi.atm_t.z.offset == 8 + 256 * 8
    assert abi.obs_t.tau.offset == 8 + 10 * 256 * 8
    assert abi.obs_t.rad.offset == abi.obs_t.tau.offset + 128 * 256 * 8
    assert abi.ctl_t.nd.offset == 4 + 8 * 10000
