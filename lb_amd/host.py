"""ctypes view of the host-side setup library (include/lbm_host.h)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import abi

_lib = None


def load_host() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(abi.HOST_LIB):
            raise RuntimeError(f"host library {abi.HOST_LIB} is not built; run __graft_entry__.build()")
        abi.load_core()
        L = ctypes.CDLL(abi.HOST_LIB)
        i32p = ctypes.POINTER(ctypes.c_int32)
        f3 = ctypes.POINTER(ctypes.c_float)
        d3 = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        sig = {
            "lbmh_voxelize_off": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_float, ctypes.c_float, d3, ctypes.c_int,
                                                 i32p, f3, f3, i32p]),
            "lbmh_mesh_write": (ctypes.c_int, [ctypes.c_char_p, i32p, ctypes.c_float, ctypes.c_float, f3, f3, i32p]),
            "lbmh_mesh_read": (ctypes.c_int, [ctypes.c_char_p, i32p, f3, f3, f3, f3, i32p]),
            "lbmh_mesher_create": (ctypes.c_int, [i32p, i32p, ctypes.c_float, ctypes.c_float, f3,
                                                  ctypes.POINTER(vp)]),
            "lbmh_mesher_wall": (ctypes.c_int, [vp]),
            "lbmh_mesher_inlet": (ctypes.c_int, [vp, d3, d3, d3, ctypes.c_int32, f3, f3, ctypes.c_int32]),
            "lbmh_mesher_inlet_paraboloid_circle": (ctypes.c_int, [vp, d3, d3, d3, ctypes.c_int32, f3, d3,
                                                                   ctypes.c_float, ctypes.c_float, ctypes.c_int32]),
            "lbmh_mesher_inlet_paraboloid_rectangle": (ctypes.c_int, [vp, d3, d3, d3, ctypes.c_int32, f3, d3, f3,
                                                                      ctypes.c_float, ctypes.c_int32]),
            "lbmh_mesher_outlet": (ctypes.c_int, [vp, d3, d3, d3, ctypes.c_int32, f3, ctypes.c_float]),
            "lbmh_mesher_periodic": (ctypes.c_int, [vp, d3, d3, d3, ctypes.c_int32, f3, ctypes.c_int32]),
            "lbmh_mesher_compile": (ctypes.c_int, [vp, i32p, ctypes.POINTER(abi.BcNode), i32p]),
            "lbmh_mesher_destroy": (None, [vp]),
            "lbmh_case_create": (ctypes.c_int, [ctypes.POINTER(vp)]),
            "lbmh_case_set": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_double]),
            "lbmh_case_get": (ctypes.c_int, [vp, ctypes.c_char_p, d3]),
            "lbmh_case_set_results_dir": (ctypes.c_int, [vp, ctypes.c_char_p]),
            "lbmh_case_load_mesh": (ctypes.c_int, [vp, ctypes.c_char_p]),
            "lbmh_case_set_grid": (ctypes.c_int, [vp, i32p, i32p, ctypes.c_float, ctypes.c_float, f3, f3]),
            "lbmh_case_mesh_wall": (ctypes.c_int, [vp]),
            "lbmh_case_inlet": (ctypes.c_int, [vp, d3, d3, d3, ctypes.c_int32, f3, f3, ctypes.c_int32]),
            "lbmh_case_inlet_paraboloid_circle": (ctypes.c_int, [vp, d3, d3, d3, ctypes.c_int32, f3, d3,
                                                                 ctypes.c_float, ctypes.c_float, ctypes.c_int32]),
            "lbmh_case_inlet_paraboloid_rectangle": (ctypes.c_int, [vp, d3, d3, d3, ctypes.c_int32, f3, d3, f3,
                                                                    ctypes.c_float, ctypes.c_int32]),
            "lbmh_case_outlet": (ctypes.c_int, [vp, d3, d3, d3, ctypes.c_int32, f3, ctypes.c_float]),
            "lbmh_case_periodic": (ctypes.c_int, [vp, d3, d3, d3, ctypes.c_int32, f3, ctypes.c_int32]),
            "lbmh_case_compile": (ctypes.c_int, [vp]),
            "lbmh_case_convert_to_lattice": (ctypes.c_int, [vp]),
            "lbmh_case_initialize": (ctypes.c_int, [vp]),
            "lbmh_case_run": (ctypes.c_int, [vp]),
            "lbmh_case_flags": (ctypes.c_int, [vp, i32p]),
            "lbmh_case_lattice": (ctypes.c_void_p, [vp]),
            "lbmh_case_errors": (ctypes.c_int, [vp, d3, i32p]),
            "lbmh_case_time_avg": (ctypes.c_int, [vp, d3, d3, d3, d3, ctypes.POINTER(ctypes.c_int64)]),
            "lbmh_read_vtk_fields": (ctypes.c_int, [ctypes.c_char_p, i32p, d3, d3, d3, d3]),
            "lbmh_case_set_init_file": (ctypes.c_int, [vp, ctypes.c_char_p]),
            "lbmh_case_write_vtk": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_int32]),
            "lbmh_case_destroy": (None, [vp]),
        }
        for k, (r, a) in sig.items():
            fn = getattr(L, k)
            fn.restype, fn.argtypes = r, a
        _lib = L
    return _lib


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def _d3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


def _i32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def voxelize_off(path, resolution, additional, outside, threads=0, device=False):
    """0/1 grid shaped (nz, ny, nx) plus (bbmin, bbmax) as the reference preprocess computes them.
    device=True runs the inside tests on the GPU (same grid)."""
    L = load_host()
    if device:
        threads = -1
    dims = (ctypes.c_int32 * 3)()
    mn, mx = (ctypes.c_float * 3)(), (ctypes.c_float * 3)()
    st = L.lbmh_voxelize_off(path.encode(), resolution, additional, _d3(outside), threads, dims, mn, mx, None)
    if st != 0:
        raise RuntimeError(f"voxelize_off({path}): {'empty polyhedron' if st == 1 else 'unreadable OFF'} ({st})")
    g = np.zeros(dims[2] * dims[1] * dims[0], np.int32)
    st = L.lbmh_voxelize_off(path.encode(), resolution, additional, _d3(outside), threads, dims, mn, mx, _i32(g))
    if st != 0:
        raise RuntimeError(f"voxelize_off failed ({st})")
    return g.reshape(dims[2], dims[1], dims[0]), tuple(mn), tuple(mx)


def mesh_write(path, grid, resolution, additional, bbmin, bbmax):
    g = np.ascontiguousarray(grid, dtype=np.int32)
    nz, ny, nx = g.shape
    dims = (ctypes.c_int32 * 3)(nx, ny, nz)
    if load_host().lbmh_mesh_write(path.encode(), dims, resolution, additional, _f3(bbmin), _f3(bbmax), _i32(g)):
        raise RuntimeError(f"mesh_write({path}) failed")


def mesh_read(path):
    L = load_host()
    dims = (ctypes.c_int32 * 3)()
    res, add = ctypes.c_float(), ctypes.c_float()
    mn, mx = (ctypes.c_float * 3)(), (ctypes.c_float * 3)()
    if L.lbmh_mesh_read(path.encode(), dims, ctypes.byref(res), ctypes.byref(add), mn, mx, None):
        raise RuntimeError(f"mesh_read({path}): unreadable or malformed")
    g = np.zeros(dims[0] * dims[1] * dims[2], np.int32)
    if L.lbmh_mesh_read(path.encode(), dims, ctypes.byref(res), ctypes.byref(add), mn, mx, _i32(g)):
        raise RuntimeError(f"mesh_read({path}): truncated or malformed")
    return g.reshape(dims[2], dims[1], dims[0]), res.value, add.value, tuple(mn), tuple(mx)


class Mesher:
    """geometryMesherBoundary over a 0/1 grid (nz, ny, nx)."""

    def __init__(self, grid01, resolution=1.0, additional=1.0, bbmin=(0.0, 0.0, 0.0)):
        self._L = load_host()
        g = np.ascontiguousarray(grid01, dtype=np.int32)
        self.shape = g.shape
        nz, ny, nx = g.shape
        dims = (ctypes.c_int32 * 3)(nx, ny, nz)
        h = ctypes.c_void_p()
        if self._L.lbmh_mesher_create(dims, _i32(g), resolution, additional, _f3(bbmin), ctypes.byref(h)):
            raise RuntimeError("lbmh_mesher_create failed")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.lbmh_mesher_destroy(self._h)
            self._h = None

    def wall(self):
        self._L.lbmh_mesher_wall(self._h)

    def inlet(self, p1, p2, p3, cid, normal, velocity, kind=abi.BC_INLET_EQUILIBRIUM):
        return self._L.lbmh_mesher_inlet(self._h, _d3(p1), _d3(p2), _d3(p3), cid, _f3(normal), _f3(velocity), kind)

    def inlet_paraboloid_circle(self, p1, p2, p3, cid, normal, center, radius, mean_velocity,
                                kind=abi.BC_INLET_EQUILIBRIUM):
        return self._L.lbmh_mesher_inlet_paraboloid_circle(self._h, _d3(p1), _d3(p2), _d3(p3), cid, _f3(normal),
                                                           _d3(center), radius, mean_velocity, kind)

    def inlet_paraboloid_rectangle(self, p1, p2, p3, cid, normal, center, dims, mean_velocity,
                                   kind=abi.BC_INLET_EQUILIBRIUM):
        """Returns the inlet extents after the call (the one along the normal becomes 1)."""
        d = (ctypes.c_float * 3)(*dims)
        st = self._L.lbmh_mesher_inlet_paraboloid_rectangle(self._h, _d3(p1), _d3(p2), _d3(p3), cid, _f3(normal),
                                                            _d3(center), d, mean_velocity, kind)
        if st:
            raise RuntimeError("lbmh_mesher_inlet_paraboloid_rectangle failed")
        return tuple(d)

    def outlet(self, p1, p2, p3, cid, normal, density):
        return self._L.lbmh_mesher_outlet(self._h, _d3(p1), _d3(p2), _d3(p3), cid, _f3(normal), density)

    def periodic(self, p1, p2, p3, cid, normal, per_idx):
        return self._L.lbmh_mesher_periodic(self._h, _d3(p1), _d3(p2), _d3(p3), cid, _f3(normal), per_idx)

    def compile(self):
        flags = np.zeros(self.shape, np.int32)
        n = ctypes.c_int32(0)
        self._L.lbmh_mesher_compile(self._h, None, None, ctypes.byref(n))
        arr = (abi.BcNode * max(n.value, 1))()
        cap = ctypes.c_int32(n.value)
        if self._L.lbmh_mesher_compile(self._h, _i32(flags), arr, ctypes.byref(cap)):
            raise RuntimeError("lbmh_mesher_compile failed")
        return flags, list(arr)[: cap.value]
