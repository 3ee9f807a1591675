"""Output path: asynchronous snapshots (lbm_snapshot_begin / _wait), the
single-plane population read, and the outputerVTK variants of the host
driver (variablesTimeAvgLatticeVTK, distributionFunctionVTK, tauVTK;
postprocesors/outputerVTK.h:449-822), against the synchronous reads."""
import ctypes
import os

import numpy as np
import pytest

from helpers import bc_nodes
from lb_amd import Lattice, LbmError, abi, host, make_desc
from oracle import mesher_oracle as mo

pytestmark = pytest.mark.gpu


def _channel_lattice(storage=abi.STORAGE_F32):
    flags, bcd = mo.channel(40, 20, 16)
    L = Lattice(make_desc(flags, bc_nodes(bcd), collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY,
                          c_les=0.17, storage=storage, arith=abi.ARITH_FAST, compat=abi.COMPAT_REFERENCE, tau=0.53,
                          nu=0.01))
    rng = np.random.default_rng(7)
    solid = flags == 0
    L.init_equilibrium(None, np.where(solid, 0.0, 0.04 + 0.004 * rng.standard_normal(flags.shape)), None, None)
    return L


def test_snapshot_is_the_state_it_was_taken_from(gpu):
    a, b = _channel_lattice(), _channel_lattice()
    for L in (a, b):
        L.step(10)
    a.snapshot_begin()
    with pytest.raises(LbmError):
        a.snapshot_begin()          # one snapshot at a time
    a.step(7)                       # queued behind the snapshot, runs while it is copied
    step, snap = a.snapshot_wait()
    assert step == 10 and a.steps() == 17
    ref = b.macroscopic() + [b.omega()]
    for k in range(5):
        assert np.array_equal(snap[k], ref[k]), k
    with pytest.raises(LbmError):
        a.snapshot_wait()           # nothing pending
    # the stepping went on: the lattice stepped 7 more times is the same as b stepped 7 more
    b.step(7)
    assert np.array_equal(a.macroscopic()[1], b.macroscopic()[1])


def test_population_plane_equals_full_read(gpu):
    L = _channel_lattice(abi.STORAGE_F16DEV)
    L.step(5)
    f = L.populations()
    for q in (0, 5, 20, 26):
        assert np.array_equal(L.population(q), f[q])


def _cavity_case(tmp_path, n=12, steps=60, **extra):
    Lh = host.load_host()
    h = ctypes.c_void_p()
    assert Lh.lbmh_case_create(ctypes.byref(h)) == 0
    grid = np.zeros((n + 2,) * 3, np.int32)
    grid[1:-1, 1:-1, 1:-1] = 1
    dims = (ctypes.c_int32 * 3)(n + 2, n + 2, n + 2)
    assert Lh.lbmh_case_set_grid(h, dims, grid.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 1.0, 1.0,
                                 (ctypes.c_float * 3)(0, 0, 0), (ctypes.c_float * 3)(n, n, n)) == 0
    tau = 0.56
    params = dict(tau=tau, ny=(tau - 0.5) / 3, rho_fyz=1.0, Cm=1.0, Cl=1.0, Cu_inverse=1.0, iterations=steps,
                  err_every_it=-1, plot_every_it=-1, iterationsMomentAvg=10, collision=abi.COLLISION_CUM2015,
                  closure=abi.CLOSURE_SMAGORINSKY, CLES=0.17, storage=abi.STORAGE_F32, arith=abi.ARITH_FAST,
                  compat=abi.COMPAT_REFERENCE)
    params.update(extra)
    for k, v in params.items():
        assert Lh.lbmh_case_set(h, k.encode(), float(v)) == 0, k
    os.makedirs(tmp_path, exist_ok=True)
    assert Lh.lbmh_case_set_results_dir(h, (str(tmp_path) + "/").encode()) == 0
    assert Lh.lbmh_case_mesh_wall(h) == 0
    lid = (ctypes.c_float * 3)(float(np.float32(0.05)), 0.0, 0.0)
    assert Lh.lbmh_case_inlet(h, (ctypes.c_double * 3)(-1.0, n - 0.75, -1.0),
                              (ctypes.c_double * 3)(n + 1.0, n - 0.25, n + 1.0),
                              (ctypes.c_double * 3)(-1.0, n - 0.75, n + 1.0), -1, (ctypes.c_float * 3)(0, 1, 0),
                              lid, abi.BC_INLET_MOVING_WALL) == 0
    assert Lh.lbmh_case_compile(h) == 0
    assert Lh.lbmh_case_initialize(h) == 0
    return Lh, h


def _read_vtk_scalars(path):
    """The numbers after each LOOKUP_TABLE / VECTORS line, as float arrays
    (inf and nan included: 1/omega on solid cells is inf)."""
    blocks, cur = [], None
    with open(path) as fh:
        for line in fh:
            if line.startswith("LOOKUP_TABLE") or line.startswith("VECTORS"):
                cur = []
                blocks.append(cur)
            elif line.startswith("SCALARS"):
                cur = None
            elif cur is not None and line.strip():
                cur.extend(float(v) for v in line.split())
    return [np.array(b) for b in blocks]


def test_async_output_writes_the_same_files(gpu, tmp_path):
    files = {}
    for mode in (0, 1):
        d = tmp_path / f"mode{mode}"
        Lh, h = _cavity_case(d, plot_every_it=10, async_output=mode)
        assert Lh.lbmh_case_run(h) == 0
        Lh.lbmh_case_destroy(h)
        files[mode] = {f: open(d / f, "rb").read() for f in sorted(os.listdir(d)) if f.startswith("variables.")}
    assert len(files[0]) == 5 and files[0].keys() == files[1].keys()   # plots at steps 10 .. 50
    for f in files[0]:
        assert files[0][f] == files[1][f], f


def test_vtk_variants_write_what_the_reads_return(gpu, tmp_path):
    Lh, h = _cavity_case(tmp_path)
    assert Lh.lbmh_case_run(h) == 0
    ctx = Lh.lbmh_case_lattice(h)
    lib = abi.load_core()
    shape = (14, 14, 14)
    om = np.empty(shape)
    assert lib.lbm_get_omega(ctx, abi.ptr(om)) == 0
    f5 = np.empty(shape)
    assert lib.lbm_get_population(ctx, 5, abi.ptr(f5)) == 0
    avg = [np.empty(shape) for _ in range(4)]
    cnt = ctypes.c_int64()
    assert Lh.lbmh_case_time_avg(h, *[abi.ptr(a) for a in avg], ctypes.byref(cnt)) == 0 and cnt.value > 0
    for which, idx in (("tau", 0), ("distributionFunction", 5), ("variablesTimeAvgLattice", 0)):
        assert Lh.lbmh_case_write_vtk(h, which.encode(), idx) == 0, which
    assert Lh.lbmh_case_write_vtk(h, b"distributionFunction", 27) == 1
    assert Lh.lbmh_case_write_vtk(h, b"nonsense", 0) == 1
    # six significant digits in the files (the stream's default precision)
    close = lambda a, b: np.allclose(a, b, rtol=1e-5, atol=1e-12)  # noqa: E731
    (tau,) = _read_vtk_scalars(tmp_path / "tau.vtk")
    fluid = om.ravel() > 0
    assert close(tau[fluid], 1.0 / om.ravel()[fluid]) and np.all(np.isinf(tau[~fluid]))
    (df,) = _read_vtk_scalars(tmp_path / "distributionFunction5.vtk")
    assert close(df, f5.ravel())
    assert "SCALARS DFunciton_5 double 1" in open(tmp_path / "distributionFunction5.vtk").read()
    rho, p, u = _read_vtk_scalars(tmp_path / "variablesTimeAvg.vtk")
    assert close(rho, avg[0].ravel()) and close(p, avg[0].ravel() / 3)
    assert close(u.reshape(-1, 3), np.stack([a.ravel() for a in avg[1:]], axis=1))
    Lh.lbmh_case_destroy(h)
