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


def test_face_plan_counts_and_rejects_bad_input(libpath):
    """pc_amr_face_plan on a 2x1 level-1 grid (host only): counts with NULL
    outputs, the rows themselves, and PC_E_INVALID for an index outside its
    level's grid or beyond the hash key's 25 bits, an uncovered face and too
    small an output."""
    import numpy as np
    from peanoclaw_amd import _abi
    lib = _abi.load()
    vp = ctypes.c_void_p

    def plan(level, ij, leaves, px=2, py=1, cap_g=None, cap_w=None, outputs=True):
        level = np.ascontiguousarray(level, np.int32)
        ij = np.ascontiguousarray(ij, np.int32).reshape(-1)
        handle = np.arange(len(level), dtype=np.int32) + 10
        leaves = np.ascontiguousarray(leaves, np.int32)
        gh = np.zeros((8 * len(leaves) + 1, 9), np.int32)
        wl = np.zeros((4 * len(leaves) + 1, 3), np.int32)
        ng, nw = ctypes.c_int(), ctypes.c_int()
        rc = lib.pc_amr_face_plan(px, py, 4, len(level), level.ctypes.data_as(vp),
                                  ij.ctypes.data_as(vp), handle.ctypes.data_as(vp),
                                  leaves.ctypes.data_as(vp), len(leaves),
                                  gh.ctypes.data_as(vp) if outputs else None,
                                  len(gh) if cap_g is None else cap_g, ctypes.byref(ng),
                                  wl.ctypes.data_as(vp) if outputs else None,
                                  len(wl) if cap_w is None else cap_w, ctypes.byref(nw))
        return rc, gh[:ng.value], wl[:nw.value], ng.value, nw.value

    # two level-1 patches side by side: each has one face neighbour and
    # three walls; no corner lies inside the domain
    rc, gh, wl, ng, nw = plan([1, 1], [[0, 0], [1, 0]], [0, 1], outputs=False)
    assert rc == 0 and (ng, nw) == (2, 6)
    rc, gh, wl, ng, nw = plan([1, 1], [[0, 0], [1, 0]], [0, 1])
    assert rc == 0
    # x+ face of patch 0 from patch 1 (same level: kind 0), x- face of patch 1
    assert gh[0].tolist() == [0, 11, 10, 1, 4, 0, 0, 4, 0]
    assert gh[1].tolist() == [0, 10, 11, 1, 4, 3, 0, -1, 0]
    # walls left, top, right, bottom as (slot, dimension, upper)
    assert wl[:3].tolist() == [[10, 0, 0], [10, 1, 1], [10, 1, 0]]
    assert wl[3:].tolist() == [[11, 1, 1], [11, 0, 1], [11, 1, 0]]
    assert plan([1, 1], [[0, 0], [1 << 25, 0]], [0])[0] == _abi.PC_E_INVALID
    assert plan([1, 13], [[0, 0], [1, 0]], [0])[0] == _abi.PC_E_INVALID
    assert plan([1, 1], [[0, 0], [2, 0]], [0])[0] == _abi.PC_E_INVALID   # outside the grid
    assert plan([1], [[0, 0]], [0])[0] == _abi.PC_E_INVALID      # x+ face uncovered
    assert plan([1, 1], [[0, 0], [1, 0]], [0, 1], cap_g=1)[0] == _abi.PC_E_INVALID
    assert plan([1, 1], [[0, 0], [1, 0]], [0, 1], cap_w=5)[0] == _abi.PC_E_INVALID
