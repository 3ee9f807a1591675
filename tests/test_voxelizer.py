"""OFF voxelization: the oracle and the product against the reference's
committed grid (3D/simulation/results/BackwardStepLaminar.txt, produced by the
reference preprocess from 3D/preprocess/BackwardStepLaminar.off with
resolution 0.25, additional factor 1)."""
import ctypes
import gzip
import os

import numpy as np

from helpers import ROOT, VOXEL_LIB

GOLD_OFF = os.path.join(ROOT, "tests", "golden", "BackwardStepLaminar.off")
GOLD_TXT = os.path.join(ROOT, "tests", "golden", "BackwardStepLaminar.txt.gz")
OUTSIDE = (-3.1416, -2.7183, -1.4142)


def golden_grid():
    lines = gzip.open(GOLD_TXT, "rt").read().split("\n")
    dims = [int(v) for v in lines[0].split()[1:4]]
    vals = np.array(" ".join(lines[9:]).split(), dtype=np.int32)
    return dims, vals.reshape(dims[2], dims[1], dims[0])


def oracle_voxelize(path, res, add, outside):
    L = ctypes.CDLL(VOXEL_LIB)
    V = ctypes.POINTER(ctypes.c_double)()
    T = ctypes.POINTER(ctypes.c_int)()
    nv, nt = ctypes.c_int(), ctypes.c_int()
    assert L.vox_read_off(path.encode(), ctypes.byref(V), ctypes.byref(nv), ctypes.byref(T), ctypes.byref(nt)) == 0
    mn, mx = (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
    L.vox_bbox(V, nv, mn, mx)
    fmn, fmx = (ctypes.c_float * 3)(*mn), (ctypes.c_float * 3)(*mx)
    dims = (ctypes.c_int * 3)()
    L.vox_dims(fmn, fmx, ctypes.c_float(res), ctypes.c_float(add), dims)
    g = np.zeros(dims[0] * dims[1] * dims[2], np.int32)
    L.vox_voxelize(V, T, nt, fmn, ctypes.c_float(res), ctypes.c_float(add), dims, (ctypes.c_double * 3)(*outside),
                   g.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    L.vox_free(V)
    L.vox_free(T)
    return list(dims), g.reshape(dims[2], dims[1], dims[0])


def test_oracle_matches_reference_voxelization():
    dims, gold = golden_grid()
    assert dims == [377, 52, 52]
    d, g = oracle_voxelize(GOLD_OFF, 0.25, 1.0, OUTSIDE)
    assert d == dims
    assert np.array_equal(g, gold)
    # the step geometry: 375*50*50 box minus a 75*50*25 step
    assert int(gold.sum()) == 375 * 50 * 50 - 75 * 50 * 25


def test_oracle_empty_and_malformed_off(tmp_path):
    L = ctypes.CDLL(VOXEL_LIB)
    V = ctypes.POINTER(ctypes.c_double)()
    T = ctypes.POINTER(ctypes.c_int)()
    nv, nt = ctypes.c_int(), ctypes.c_int()
    p = tmp_path / "empty.off"
    p.write_text("OFF\n0 0 0\n")
    assert L.vox_read_off(str(p).encode(), ctypes.byref(V), ctypes.byref(nv), ctypes.byref(T), ctypes.byref(nt)) == 1
    p.write_text("OFF\n3 1 0\n0 0 0\n1 0 0\n")
    assert L.vox_read_off(str(p).encode(), ctypes.byref(V), ctypes.byref(nv), ctypes.byref(T), ctypes.byref(nt)) == 2
    assert L.vox_read_off(b"/nonexistent.off", ctypes.byref(V), ctypes.byref(nv), ctypes.byref(T),
                          ctypes.byref(nt)) == 2
