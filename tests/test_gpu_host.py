"""The C++ host driver (include/lbm_host.h) end to end on the GPU: a lid-driven
cavity set up through the case API (grid, wall meshing, moving-wall lid,
compile, initialise, run) against the oracle run from the numpy mesher's
flags and lists (the two meshers agree, tests/test_host.py)."""
import ctypes

import numpy as np
import pytest

from helpers import OracleLattice, cavity_setup
from lb_amd import abi, host, make_desc

pytestmark = pytest.mark.gpu


def test_case_cavity_run_matches_oracle(gpu, tmp_path):
    n, steps, tau = 10, 60, 0.56
    L = host.load_host()
    h = ctypes.c_void_p()
    assert L.lbmh_case_create(ctypes.byref(h)) == 0
    try:
        grid = np.zeros((n + 2, n + 2, n + 2), np.int32)
        grid[1:-1, 1:-1, 1:-1] = 1
        dims = (ctypes.c_int32 * 3)(n + 2, n + 2, n + 2)
        assert L.lbmh_case_set_grid(h, dims, grid.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 1.0, 1.0,
                                    (ctypes.c_float * 3)(0, 0, 0), (ctypes.c_float * 3)(n, n, n)) == 0
        for k, v in (("tau", tau), ("ny", (tau - 0.5) / 3), ("rho_fyz", 1.0), ("Cm", 1.0), ("Cl", 1.0),
                     ("Cu_inverse", 1.0), ("iterations", steps), ("err_every_it", 20), ("plot_every_it", -1),
                     ("iterationsMomentAvg", 10), ("probe_iterations", 30), ("ProbeLocation_x", 5),
                     ("ProbeLocation_y", 7), ("ProbeLocation_z", 4),
                     ("probe_every_it", 5), ("collision", abi.COLLISION_CUM2015), ("closure", abi.CLOSURE_NONE),
                     ("storage", abi.STORAGE_F64), ("arith", abi.ARITH_LITERAL), ("compat", abi.COMPAT_REFERENCE)):
            assert L.lbmh_case_set(h, k.encode(), float(v)) == 0, k
        assert L.lbmh_case_set_results_dir(h, str(tmp_path).encode()) == 0
        assert L.lbmh_case_mesh_wall(h) == 0
        lid = (ctypes.c_float * 3)(float(np.float32(0.05)), 0.0, 0.0)
        assert L.lbmh_case_inlet(h, (ctypes.c_double * 3)(-1.0, n - 0.75, -1.0),
                                 (ctypes.c_double * 3)(n + 1.0, n - 0.25, n + 1.0),
                                 (ctypes.c_double * 3)(-1.0, n - 0.75, n + 1.0), -1, (ctypes.c_float * 3)(0, 1, 0),
                                 lid, abi.BC_INLET_MOVING_WALL) == 0
        assert L.lbmh_case_compile(h) == 0
        assert L.lbmh_case_initialize(h) == 0
        assert L.lbmh_case_run(h) == 0
        flags = np.zeros((n + 2,) * 3, np.int32)
        assert L.lbmh_case_flags(h, flags.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))) == 0
        ctx = L.lbmh_case_lattice(h)
        lib = abi.load_core()
        out = [np.empty((n + 2,) * 3) for _ in range(4)]
        assert lib.lbm_get_macroscopic(ctx, *[abi.ptr(a) for a in out]) == 0
        errs = (ctypes.c_double * 8)()
        ne = ctypes.c_int32(8)
        assert L.lbmh_case_errors(h, errs, ctypes.byref(ne)) == 0
        units = {}
        for k in ("Cu", "Crho", "Cpressure", "cs2"):
            v = ctypes.c_double()
            assert L.lbmh_case_get(h, k.encode(), ctypes.byref(v)) == 0
            units[k] = v.value
        avg = [np.empty((n + 2,) * 3) for _ in range(4)]
        cnt = ctypes.c_int64()
        assert L.lbmh_case_time_avg(h, *[abi.ptr(a) for a in avg], ctypes.byref(cnt)) == 0
    finally:
        L.lbmh_case_destroy(h)
    oflags, bc, ux = cavity_setup(n)
    assert np.array_equal(flags, oflags)
    # LBMConstants fields are float (RealType), so the case runs with float(tau)
    o = OracleLattice(make_desc(oflags, bc, collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F64,
                                arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE, tau=float(np.float32(tau))))
    o.init_equilibrium(None, ux, None, None)
    # the reference loop evaluates the error at the end of iteration k with
    # k % err_every_it == 0, k != 0, i.e. after k + 1 steps (SolverTurbulentLES.h:184-251)
    ref_err = []
    o.step(21)
    ref_err.append(o.error_quadratic())
    o.step(20)
    ref_err.append(o.error_quadratic())
    # momentAdd after the steps k > iterations - iterationsMomentAvg = 50
    o.step(10)
    o.step_moment_add(steps - 51)
    mo = o.macroscopic()
    for a, b in zip(out, mo):
        assert np.array_equal(a, b)
    # probe rows at k = 30, 35, ..., 55 (k % probe_every_it == 0, k >= iterations -
    # probe_iterations), written from device records after the run
    rows = (tmp_path / "probeData.csv").read_text().splitlines()
    assert rows[1] == "ux,uy,uz,rho,pressure"
    vals = np.array([[float(t) for t in r.split(",")] for r in rows[2:]])
    op = OracleLattice(make_desc(oflags, bc, collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F64,
                                 arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE, tau=float(np.float32(tau))))
    op.init_equilibrium(None, ux, None, None)
    want = []
    done = 0
    for k in range(30, 60, 5):
        op.step(k + 1 - done)
        done = k + 1
        m = op.macroscopic()
        r, vx, vy, vz = (float(np.float32(m[j][4, 7, 5])) for j in range(4))
        f32 = np.float32
        want.append([f32(f32(vx) * f32(units["Cu"])), f32(f32(vy) * f32(units["Cu"])), f32(f32(vz) * f32(units["Cu"])),
                     f32(f32(r) * f32(units["Crho"])), f32(f32(r) * f32(units["cs2"]) * f32(units["Cpressure"]))])
    print(f"[host case] probe rows {len(vals)}", flush=True)
    assert vals.shape == (6, 5)
    assert np.allclose(vals, np.array(want, dtype=np.float64), rtol=2e-5, atol=1e-9)
    oavg, ocnt = o.moment_avg()
    assert cnt.value == ocnt == 9
    fluid = oflags != 0
    for a, b in zip(avg, oavg):
        assert np.array_equal(a[fluid], b[fluid])
    got = [errs[i] for i in range(ne.value)]
    print(f"[host case] errors {got} oracle {ref_err}", flush=True)
    assert len(got) >= 2
    assert np.allclose(got, ref_err[:len(got)], rtol=1e-6)


def test_case_initialized_from_vtk_fields_matches_oracle(gpu, tmp_path):
    """InitializationEquilibriumVariables through the case API: the fields file
    (lattice units) sets rho and u, inlet cells take the lid velocity, and the
    run equals the oracle started from the same (float-rounded) fields."""
    from test_host import _vtk_fields_text
    n, steps, tau = 8, 20, 0.56
    dims = (n + 2, n + 2, n + 2)
    rng = np.random.default_rng(31)
    rho = (1 + 0.01 * rng.standard_normal(dims[::-1])).astype(np.float32).astype(np.float64)
    u = [(0.02 * rng.standard_normal(dims[::-1])).astype(np.float32).astype(np.float64) for _ in range(3)]
    (tmp_path / "init.vtk").write_text(_vtk_fields_text(dims, rho, u))
    L = host.load_host()
    h = ctypes.c_void_p()
    assert L.lbmh_case_create(ctypes.byref(h)) == 0
    try:
        grid = np.zeros(dims[::-1], np.int32)
        grid[1:-1, 1:-1, 1:-1] = 1
        assert L.lbmh_case_set_grid(h, (ctypes.c_int32 * 3)(*dims), grid.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                    1.0, 1.0, (ctypes.c_float * 3)(0, 0, 0), (ctypes.c_float * 3)(n, n, n)) == 0
        for k, v in (("tau", tau), ("ny", (tau - 0.5) / 3), ("rho_fyz", 1.0), ("Cm", 1.0), ("Cl", 1.0),
                     ("Cu_inverse", 1.0), ("iterations", steps), ("err_every_it", -1), ("plot_every_it", -1),
                     ("probe_every_it", -1), ("collision", abi.COLLISION_CUM2015), ("closure", abi.CLOSURE_NONE),
                     ("storage", abi.STORAGE_F64), ("arith", abi.ARITH_LITERAL), ("compat", abi.COMPAT_REFERENCE)):
            assert L.lbmh_case_set(h, k.encode(), float(v)) == 0, k
        assert L.lbmh_case_set_results_dir(h, str(tmp_path).encode()) == 0
        assert L.lbmh_case_set_init_file(h, b"init.vtk") == 0
        assert L.lbmh_case_mesh_wall(h) == 0
        lid = (ctypes.c_float * 3)(float(np.float32(0.05)), 0.0, 0.0)
        assert L.lbmh_case_inlet(h, (ctypes.c_double * 3)(-1.0, n - 0.75, -1.0),
                                 (ctypes.c_double * 3)(n + 1.0, n - 0.25, n + 1.0),
                                 (ctypes.c_double * 3)(-1.0, n - 0.75, n + 1.0), -1, (ctypes.c_float * 3)(0, 1, 0),
                                 lid, abi.BC_INLET_MOVING_WALL) == 0
        assert L.lbmh_case_compile(h) == 0
        assert L.lbmh_case_initialize(h) == 0
        assert L.lbmh_case_run(h) == 0
        ctx = L.lbmh_case_lattice(h)
        out = [np.empty(dims[::-1]) for _ in range(4)]
        assert abi.load_core().lbm_get_macroscopic(ctx, *[abi.ptr(a) for a in out]) == 0
    finally:
        L.lbmh_case_destroy(h)
    oflags, bc, ux_lid = cavity_setup(n)
    solid = oflags == 0
    lidc = ux_lid != 0
    o = OracleLattice(make_desc(oflags, bc, collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F64,
                                arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE, tau=float(np.float32(tau))))
    o.init_equilibrium(np.where(solid, 0.0, rho), np.where(lidc, ux_lid, np.where(solid, 0.0, u[0])),
                       np.where(lidc, 0.0, np.where(solid, 0.0, u[1])), np.where(lidc, 0.0, np.where(solid, 0.0, u[2])))
    o.step(steps)
    fluid = ~solid
    for a, b in zip(out, o.macroscopic()):
        assert np.array_equal(a[fluid], b[fluid])
