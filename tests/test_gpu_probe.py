"""Probe records without synchronisation (lbm_probe_record / lbm_probe_fetch)
against the synchronous probe and the full read-back, through the C ABI.

The reference's probeWrite copies four values to the host and waits every
probe step (SolverTurbulentLES.h:150-177); the records here are made on the
device and read in one copy.  They must equal lbm_probe and
lbm_get_macroscopic at the probe cell bit for bit, keep their order across
the device buffer's refills, and follow the fetch contract.
"""
import numpy as np
import pytest

from helpers import random_velocity_field
from lb_amd import Lattice, LbmError, abi, make_desc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage,arith", [(abi.STORAGE_F64, abi.ARITH_LITERAL), (abi.STORAGE_F32, abi.ARITH_FAST),
                                           (abi.STORAGE_F16DEV_SR, abi.ARITH_FAST)],
                         ids=["f64_literal", "f32_fast", "f16dev_sr"])
def test_records_equal_probe_and_readback(gpu, storage, arith):
    n = 20
    flags = np.ones((n, n, n), np.int32)
    flags[8:12, 8:12, 8:12] = 0
    g = Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY, c_les=0.17,
                          storage=storage, arith=arith, compat=abi.COMPAT_REFERENCE, periodic=(1, 1, 1), tau=0.52))
    u = random_velocity_field(flags.shape, 0.04, seed=51)
    g.init_equilibrium(None, *[np.where(flags == 0, 0.0, c) for c in u])
    cells = [(3, 4, 5), (13, 7, 2), (9, 9, 9)]   # the last one is solid
    ref = []
    g.probe_record(*cells[0])                    # the initial state
    ref.append(g.probe(*cells[0]))
    for k in range(5):
        g.step(1 + k)
        for c in cells:
            g.probe_record(*c)
        m = g.macroscopic()
        for c in cells:
            p = g.probe(*c)
            assert np.array_equal(p, [m[j][c[2], c[1], c[0]] for j in range(4)])
            ref.append(p)
    rec = g.probe_fetch()
    print(f"[probe records storage={storage}] {len(rec)} records", flush=True)
    assert rec.shape == (16, 4)
    assert np.array_equal(rec, np.array(ref))
    assert np.all(rec[3::3] == 0.0)   # solid cell
    assert g.probe_fetch().shape == (0, 4)


def test_records_survive_buffer_refills_in_order(gpu):
    n = 16
    flags = np.ones((n, n, n), np.int32)
    g = Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F32,
                          arith=abi.ARITH_FAST, periodic=(1, 1, 1), tau=0.6))
    u = random_velocity_field(flags.shape, 0.03, seed=52)
    g.init_equilibrium(None, *u)
    m = g.macroscopic()
    total = 9000   # more than two device buffers
    for i in range(total):
        g.probe_record(i % n, (i // n) % n, (i // (n * n)) % n)
    rec = g.probe_fetch()
    assert rec.shape == (total, 4)
    i = np.arange(total)
    x, y, z = i % n, (i // n) % n, (i // (n * n)) % n
    for j in range(4):
        assert np.array_equal(rec[:, j], m[j][z, y, x])


def test_fetch_contract(gpu):
    n = 8
    flags = np.ones((n, n, n), np.int32)
    g = Lattice(make_desc(flags, [], collision=abi.COLLISION_SRT, storage=abi.STORAGE_F64,
                          arith=abi.ARITH_LITERAL, periodic=(1, 1, 1), tau=0.6))
    g.init_equilibrium(None, 0.01, 0.0, 0.0)
    with pytest.raises(LbmError):
        g.probe_record(n, 0, 0)
    for _ in range(3):
        g.probe_record(1, 2, 3)
    import ctypes
    cnt = ctypes.c_int64()
    lib = abi.load_core()
    assert lib.lbm_probe_fetch(g._h, None, 0, ctypes.byref(cnt)) == abi.LBM_OK and cnt.value == 3
    small = np.empty((2, 4))
    assert lib.lbm_probe_fetch(g._h, abi.ptr(small), 2, ctypes.byref(cnt)) == abi.LBM_E_ARG
    rec = g.probe_fetch()   # nothing was cleared by the failed call
    assert rec.shape == (3, 4)
    assert np.allclose(rec[:, 1], 0.01) and np.allclose(rec[:, 0], 1.0)
