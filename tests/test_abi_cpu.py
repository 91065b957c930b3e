"""CPU checks of the C-ABI library: it loads and exports every entry point
that include/peanoclaw_amd.h declares (no compute without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "peanoclaw_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pc_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def libpath():
    from peanoclaw_amd import _abi
    if not os.path.exists(_abi.LIB_PATH):
        from peanoclaw_amd import build
        build.build()
    return _abi.LIB_PATH


def test_library_exports_every_declared_symbol(libpath):
    out = subprocess.run(["nm", "-D", "--defined-only", libpath], capture_output=True,
                         text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    ctypes.CDLL(libpath)


def test_python_binding_covers_the_header(libpath):
    from peanoclaw_amd import _abi
    assert sorted(_abi.EXPORTED_SYMBOLS) == declared_functions()
    _abi.load()


def test_constants_and_no_device_error(libpath):
    from peanoclaw_amd import _abi
    lib = _abi.load()
    assert lib.pc_get_number_of_unknowns_per_cell() == 3
    assert lib.pc_get_number_of_parameter_fields_without_ghostlayer() == 0
    assert lib.pc_get_number_of_parameter_fields_with_ghostlayer() == 1
    assert lib.pc_get_ghostlayer_width() == 1
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    rc = lib.pc_pool_create(0, 0, ctypes.byref(h))
    assert rc == _abi.PC_E_NODEVICE
    assert b"device" in lib.pc_last_error_string()


def test_tile_stride_rule(libpath):
    """Rows of ny + 2 doubles below ny = 8; from 8 on, padded to a multiple
    of 16 doubles (so that the interior can start on a 128-byte line)."""
    from peanoclaw_amd import _abi
    lib = _abi.load()
    for ny in range(2, 65):
        st = lib.pc_tile_stride(ny)
        if ny < 8:
            assert st == ny + 2
        else:
            assert st % 16 == 0 and ny + 2 <= st < ny + 18
    assert lib.pc_tile_stride(1) == -1 and lib.pc_tile_stride(65) == -1
