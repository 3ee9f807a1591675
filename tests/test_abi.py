"""The C-ABI libraries load and export every declared symbol; descriptor
validation runs on the host (no GPU needed for these)."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import ROOT, cavity_setup
from lb_amd import Lattice, LbmError, abi, make_desc


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:lbm|vox|lbmh)_[a-z0-9_]+)\s*\(", text)))


def test_core_exports_every_declared_symbol():
    lib = ctypes.CDLL(abi.CORE_LIB)
    names = declared("lbm_core.h")
    assert "lbm_create" in names and "lbm_step" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert sorted(n for n, _, _ in abi.CORE_SYMBOLS) == names


def test_core_exports_every_closure_symbol():
    lib = ctypes.CDLL(abi.CORE_LIB)
    names = declared("lbm_closure.h")
    assert "lbm_mlp_train" in names and "lbm_set_closure" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(n for n, _, _ in abi.CLOSURE_SYMBOLS) == names


def test_status_strings():
    lib = abi.load_core()
    for s in range(7):
        assert lib.lbm_status_str(s)


def _create_status(desc):
    lib = abi.load_core()
    h = ctypes.c_void_p()
    st = lib.lbm_create(ctypes.byref(desc), ctypes.byref(h))
    if st == abi.LBM_OK:
        lib.lbm_destroy(h)
    return st


def test_validation_rejects_bad_descriptors():
    flags, bc, _ = cavity_setup(4)
    # tau <= 1/2 (omega >= 2)
    assert _create_status(make_desc(flags, bc, tau=0.5)) == abi.LBM_E_ARG
    # wall node whose neighbour lies outside the lattice on a non-periodic axis
    # (the reference reads outside its arrays there)
    ones = np.ones((4, 4, 4), np.int32)
    assert _create_status(make_desc(ones, [abi.BcNode(0, 1, 1, abi.BC_WALL)], tau=0.6)) == abi.LBM_E_UNSUPPORTED
    # two inlet entries on one cell
    f2 = flags.copy()
    f2[2, 2, 2] = -1
    inl = abi.BcNode(2, 2, 2, abi.BC_INLET_MOVING_WALL)
    assert _create_status(make_desc(f2, list(bc) + [inl, inl], tau=0.6)) == abi.LBM_E_ARG
    # outlet entry on an inlet cell
    out = abi.BcNode(2, 2, 2, abi.BC_OUTLET_INTERPOLATED, nx=1.0, rho=1.0)
    assert _create_status(make_desc(f2, list(bc) + [inl, out], tau=0.6)) == abi.LBM_E_UNSUPPORTED
    # periodic partner index outside the lattice
    f4 = flags.copy()
    f4[1, 1, 1] = -3
    gh = abi.BcNode(1, 1, 1, abi.BC_PERIODIC_GHOST, nx=-1.0, per_idx=4)
    assert _create_status(make_desc(f4, list(bc) + [gh], tau=0.6)) == abi.LBM_E_ARG
    # unknown negative id
    f3 = flags.copy()
    f3[2, 2, 2] = -7
    assert _create_status(make_desc(f3, bc, tau=0.6)) == abi.LBM_E_UNSUPPORTED
    # wall entry on a solid cell
    bad = list(bc) + [abi.BcNode(0, 0, 0, abi.BC_WALL)]
    assert _create_status(make_desc(flags, bad, tau=0.6)) == abi.LBM_E_ARG
    # literal arithmetic is FP64/FP32 only
    for st in (abi.STORAGE_F16DEV, abi.STORAGE_BF16DEV, abi.STORAGE_F16DEV_SR, abi.STORAGE_BF16DEV_SR):
        assert _create_status(make_desc(flags, bc, tau=0.6, storage=st, arith=abi.ARITH_LITERAL)) == abi.LBM_E_ARG
    # unknown storage code
    assert _create_status(make_desc(flags, bc, tau=0.6, storage=6, arith=abi.ARITH_FAST)) == abi.LBM_E_ARG


@pytest.mark.parametrize("storage,n_over", [
    (abi.STORAGE_F64, (1024, 1024, 512)),      # 2^29 cells x 8 B = 4 GiB per plane
    (abi.STORAGE_F32, (1024, 1024, 1024)),     # 2^30 cells x 4 B
    (abi.STORAGE_F16DEV, (2048, 1024, 1024)),  # 2^31 cells
], ids=["f64", "f32", "f16dev"])
def test_largest_lattice_limit(storage, n_over):
    """A population plane is addressed with 32-bit byte offsets: lattices whose
    padded plane reaches 4 GiB (or 2^31 cells) are refused at create, before
    the flag grid is read.  The flag buffer has the full size (zero pages the
    kernel maps lazily, so it costs no memory), so that a create which read it
    after all would still read inside it and fail this test instead of
    reading past the end."""
    nx, ny, nz = n_over
    flags = np.zeros((nz, ny, nx), np.int32)
    d = make_desc(flags, [], tau=0.6, storage=storage, arith=abi.ARITH_FAST, periodic=(1, 1, 1))
    assert (d.nx, d.ny, d.nz) == n_over
    assert _create_status(d) == abi.LBM_E_ARG


def test_valid_descriptor_without_device_reports_hip_error():
    n, _ = abi.device_info()
    if n > 0:
        pytest.skip("a GPU is visible")
    flags, bc, _ = cavity_setup(4)
    assert _create_status(make_desc(flags, bc, tau=0.6)) == abi.LBM_E_HIP
    # fluid on a non-periodic array boundary and a flag -1 cell without an
    # inlet entry are valid (the core pads a solid halo; the cell keeps its state)
    ones = np.ones((4, 4, 4), np.int32)
    ones[1, 1, 1] = -1
    assert _create_status(make_desc(ones, [], tau=0.6)) == abi.LBM_E_HIP
    with pytest.raises(LbmError):
        Lattice(make_desc(flags, bc, tau=0.6))


def test_voxelize_rejects_malformed_hierarchies_before_any_device_work():
    """lbm_voxelize checks every index the traversal will use, and that the
    hierarchy is a tree laid out parent-first (so every walk ends), on the host."""
    lib = abi.load_core()

    class Node(ctypes.Structure):
        _fields_ = [("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3), ("left", ctypes.c_int32),
                    ("right", ctypes.c_int32), ("first", ctypes.c_int32), ("count", ctypes.c_int32)]

    V = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], np.float64)
    T = np.array([0, 1, 2, 0, 1, 3, 0, 2, 3, 1, 2, 3], np.int32)
    order = np.arange(4, dtype=np.int32)
    dims = (ctypes.c_int32 * 3)(2, 2, 2)
    bb = (ctypes.c_double * 3)(0, 0, 0)
    out = (ctypes.c_double * 3)(5, 5, 5)
    grid = np.zeros(8, np.int32)
    i32 = ctypes.POINTER(ctypes.c_int32)

    def call(nodes, tris=T):
        arr = (Node * len(nodes))(*nodes)
        return lib.lbm_voxelize(abi.ptr(V), 4, tris.ctypes.data_as(i32), 4, ctypes.cast(arr, ctypes.c_void_p),
                                len(nodes), order.ctypes.data_as(i32), dims, 1.0, 1.0, bb, out,
                                grid.ctypes.data_as(i32))

    box = ((ctypes.c_double * 3)(0, 0, 0), (ctypes.c_double * 3)(1, 1, 1))
    leaf = lambda f, c: Node(box[0], box[1], -1, -1, f, c)   # noqa: E731
    # a child pointing back at its parent, shared children, a leaf past the triangles
    assert call([Node(box[0], box[1], 1, 2, 0, 0), Node(box[0], box[1], 0, 2, 0, 0), leaf(0, 4)]) == abi.LBM_E_ARG
    assert call([Node(box[0], box[1], 1, 1, 0, 0), leaf(0, 4)]) == abi.LBM_E_ARG
    assert call([Node(box[0], box[1], 1, 2, 0, 0), leaf(0, 2), Node(box[0], box[1], 1, 1, 0, 0)]) == abi.LBM_E_ARG
    assert call([leaf(2, 3)]) == abi.LBM_E_ARG
    bad = T.copy()
    bad[5] = 9
    assert call([leaf(0, 4)], bad) == abi.LBM_E_ARG
