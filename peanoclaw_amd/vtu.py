"""VTU output of patch data (the content of PeanoClaw's VTUSubgridPlotter,
grid/plotter/VTUSubgridPlotter.cpp:9-241), legacy VTK output (that of its
VTKSubgridPlotter, VTKSubgridPlotter.cpp:9-241) and readers for both.

`write_vtu` writes host arrays through the native writer (pc_vtu_write);
`write_pool_vtu` downloads slots of a device pool and writes them
(pc_pool_write_vtu); `read_vtu` parses the file back (XML header, appended
raw binary with UInt32 block sizes).
"""
import ctypes
import re

import numpy as np

from ._abi import check, load


class VtuPatch(ctypes.Structure):
    _fields_ = [
        ("subdivision_factor", ctypes.c_int32 * 2),
        ("ghostlayer_width", ctypes.c_int32),
        ("level", ctypes.c_int32),
        ("position", ctypes.c_double * 2),
        ("size", ctypes.c_double * 2),
        ("time", ctypes.c_double),
        ("timestep_size", ctypes.c_double),
        ("demanded_mesh_width", ctypes.c_double),
        ("age", ctypes.c_double),
        ("q", ctypes.POINTER(ctypes.c_double) * 3),
        ("b", ctypes.POINTER(ctypes.c_double)),
    ]


def _patch_array(patches):
    keep = []
    arr = (VtuPatch * max(len(patches), 1))()
    dp = ctypes.POINTER(ctypes.c_double)
    for k, p in enumerate(patches):
        q = np.ascontiguousarray(p["q"], dtype=np.float64)
        b = np.ascontiguousarray(p["b"], dtype=np.float64)
        nx, ny = q.shape[1], q.shape[2]
        if b.shape != (nx, ny):
            raise ValueError("b must have the shape of the patch")
        size = p["size"] if np.ndim(p["size"]) else (p["size"], p["size"])
        P = arr[k]
        P.subdivision_factor[0], P.subdivision_factor[1] = nx, ny
        P.ghostlayer_width = p.get("ghostlayer_width", 1)
        P.level = p.get("level", 1)
        P.position[0], P.position[1] = p["pos"]
        P.size[0], P.size[1] = size
        P.time = p.get("time", 0.0)
        P.timestep_size = p.get("dt", 0.0)
        P.demanded_mesh_width = p.get("demanded", 0.0)
        P.age = p.get("age", 0)
        for i in range(3):
            P.q[i] = q[i].ctypes.data_as(dp)
        P.b = b.ctypes.data_as(dp)
        keep += [q, b]
    return arr, keep


def write_vtu(path, patches):
    """patches: dicts with keys pos (x, y), size (float or (sx, sy)), q
    (array 3 x nx x ny: h, hu, hv), b (nx x ny) and optionally level, time,
    dt, demanded (width), age, ghostlayer_width."""
    arr, keep = _patch_array(patches)
    check(load().pc_vtu_write(str(path).encode(), len(patches), arr))


def write_vtk(path, patches):
    """The same patches as legacy ASCII VTK (VTKSubgridPlotter geometry:
    each patch shrunk by 1 + 0.015 * level around its centre)."""
    arr, keep = _patch_array(patches)
    check(load().pc_vtk_write(str(path).encode(), len(patches), arr))


def write_pool_vtu(pool, path, slots):
    """Current state (uNew), b and descriptor of device slots to `path`."""
    s = (ctypes.c_int32 * max(len(slots), 1))(*slots)
    check(load().pc_pool_write_vtu(pool.h, str(path).encode(), s, len(slots)))


def write_pool_vtk(pool, path, slots):
    """As write_pool_vtu, in legacy ASCII VTK."""
    s = (ctypes.c_int32 * max(len(slots), 1))(*slots)
    check(load().pc_pool_write_vtk(pool.h, str(path).encode(), s, len(slots)))


_TYPES = {"Float64": np.float64, "Float32": np.float32, "Int32": np.int32, "UInt8": np.uint8}


def read_vtu(path):
    """(points (n, 3), connectivity (m, 4), types (m,), cell arrays by name)."""
    raw = open(path, "rb").read()
    start = raw.index(b"<AppendedData")
    us = raw.index(b"_", start) + 1
    head = raw[:start].decode()
    blob = raw[us:]
    arrays = {}
    for m in re.finditer(r'<DataArray type="(\w+)" Name="(\w+)" NumberOfComponents="(\d+)" '
                         r'format="appended" offset="(\d+)"/>', head):
        typ, name, comps, off = m.group(1), m.group(2), int(m.group(3)), int(m.group(4))
        nbytes = int(np.frombuffer(blob[off:off + 4], dtype=np.uint32)[0])
        a = np.frombuffer(blob[off + 4:off + 4 + nbytes], dtype=_TYPES[typ])
        arrays[name] = a.reshape(-1, comps) if comps > 1 else a
    pts = arrays.pop("points")
    conn = arrays.pop("connectivity").reshape(-1, 4)
    offs = arrays.pop("offsets")
    types = arrays.pop("types")
    if not np.array_equal(offs, 4 * np.arange(1, len(conn) + 1)):
        raise ValueError("unexpected cell offsets")
    return pts, conn, types, arrays


def read_vtk(path):
    """(points (n, 3), connectivity (m, 4), types (m,), cell arrays by name)
    of a legacy ASCII VTK unstructured grid as write_vtk writes it."""
    tok = open(path).read().split("\n")
    if not tok[0].startswith("# vtk DataFile") or tok[2] != "ASCII" or \
            tok[3] != "DATASET UNSTRUCTURED_GRID":
        raise ValueError("not a legacy ASCII unstructured grid")
    k = 4
    nv = int(tok[k].split()[1])
    pts = np.array([[float(v) for v in tok[k + 1 + i].split()] for i in range(nv)])
    k += 1 + nv
    nc = int(tok[k].split()[1])
    cells = np.array([[int(v) for v in tok[k + 1 + i].split()] for i in range(nc)])
    if (cells[:, 0] != 4).any():
        raise ValueError("expected quadrilaterals")
    k += 1 + nc
    types = np.array([int(tok[k + 1 + i]) for i in range(nc)])
    k += 1 + nc
    if tok[k] != f"CELL_DATA {nc}":
        raise ValueError("expected cell data")
    k += 1
    arrays = {}
    while k < len(tok) and tok[k].startswith("SCALARS"):
        name = tok[k].split()[1]
        arrays[name] = np.array([float(tok[k + 2 + i]) for i in range(nc)])
        k += 2 + nc
    return pts, cells[:, 1:], types, arrays


__all__ = ["write_vtu", "write_vtk", "write_pool_vtu", "write_pool_vtk", "read_vtu", "read_vtk",
           "VtuPatch"]
