"""HIP core vs CPU oracle, through the C ABI, on identical seeded inputs.

FP64 with literal arithmetic must reproduce the oracle (the reference's AB
step restated) to <= 1e-12 relative L2 -- in practice bit for bit.  Reduced
precisions are compared against the FP64 oracle with stated tolerances.
"""
import numpy as np
import pytest

from helpers import OracleLattice, cavity_setup, random_velocity_field, rel_l2
from lb_amd import Lattice, abi, make_desc

pytestmark = pytest.mark.gpu

C1_TAU = 0.524  # nu = 0.05 * 64 / 400 = 0.008, tau = 3 nu + 1/2


def _pair(flags, bc, init, steps, chunks=None, **kw):
    """Run the GPU core and the oracle from the same equilibrium state."""
    d = make_desc(flags, bc, **kw)
    # the oracle runs in FP32 only to check the FP32 literal path; otherwise in FP64
    fp32_literal = kw.get("storage") == abi.STORAGE_F32 and kw.get("arith", abi.ARITH_LITERAL) == abi.ARITH_LITERAL
    dref = make_desc(flags, bc, **{**kw, "storage": abi.STORAGE_F32 if fp32_literal else abi.STORAGE_F64})
    g = Lattice(d)
    o = OracleLattice(dref)
    g.init_equilibrium(*init)
    o.init_equilibrium(*init)
    for n in (chunks or [steps]):
        g.step(n)
    o.step(steps)
    g.sync()
    return g, o


def _report(tag, g, o, solid):
    fg, fo = g.populations(), o.populations()
    mg, mo = g.macroscopic(), o.macroscopic()
    fluid = ~solid
    err_f = rel_l2(fg[:, fluid], fo[:, fluid])
    err_r = rel_l2(mg[0][fluid], mo[0][fluid])
    err_u = rel_l2(np.stack([m[fluid] for m in mg[1:]]), np.stack([m[fluid] for m in mo[1:]]))
    maxabs = np.abs(fg[:, fluid] - fo[:, fluid]).max()
    print(f"[{tag}] relL2 f={err_f:.3e} rho={err_r:.3e} u={err_u:.3e} max|df|={maxabs:.3e}", flush=True)
    return err_f, err_r, err_u


@pytest.mark.parametrize("coll", [abi.COLLISION_CUM2015, abi.COLLISION_SRT], ids=["cum2015", "srt"])
def test_cavity_fp64_literal_small(gpu, coll):
    flags, bc, ux = cavity_setup(8)
    g, o = _pair(flags, bc, (None, ux, None, None), 60, collision=coll, storage=abi.STORAGE_F64,
                 arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE, tau=C1_TAU)
    ef, er, eu = _report(f"cavity8 fp64 lit coll={coll}", g, o, flags == 0)
    assert ef <= 1e-12 and er <= 1e-12 and eu <= 1e-12


def test_cavity_fp64_literal_corrected_tables(gpu):
    """compat = 0: true opposite table and exact weights."""
    flags, bc, ux = cavity_setup(8)
    g, o = _pair(flags, bc, (None, ux, None, None), 40, collision=abi.COLLISION_CUM2015,
                 storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=0, tau=C1_TAU)
    ef, er, eu = _report("cavity8 fp64 lit corrected", g, o, flags == 0)
    assert ef <= 1e-12 and eu <= 1e-12


def test_chunked_steps_match(gpu):
    """Splitting the run into lbm_step calls of odd lengths changes nothing
    (A-A parity is tracked across calls)."""
    flags, bc, ux = cavity_setup(6)
    kw = dict(collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL,
              compat=abi.COMPAT_REFERENCE, tau=C1_TAU)
    a = Lattice(make_desc(flags, bc, **kw))
    b = Lattice(make_desc(flags, bc, **kw))
    for L in (a, b):
        L.init_equilibrium(None, ux, None, None)
    a.step(21)
    for n in (1, 7, 3, 10):
        b.step(n)
    assert np.array_equal(a.populations(), b.populations())


@pytest.mark.parametrize("coll", [abi.COLLISION_CUM2015, abi.COLLISION_SRT], ids=["cum2015", "srt"])
def test_periodic_les_fp64_literal(gpu, coll):
    shape = (12, 16, 20)
    flags = np.ones(shape, np.int32)
    u = random_velocity_field(shape, 0.04, seed=7)
    rho = 1.0 + 0.01 * random_velocity_field(shape, 1.0, seed=8)[0]
    g, o = _pair(flags, [], (rho, *u), 40, collision=coll, closure=abi.CLOSURE_SMAGORINSKY,
                 storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE,
                 periodic=(1, 1, 1), tau=0.505, c_les=float(np.float32(0.173)))
    ef, er, eu = _report(f"periodic les fp64 lit coll={coll}", g, o, flags == 0)
    assert ef <= 1e-12 and er <= 1e-12 and eu <= 1e-12
    wg, wo = g.omega(), o.omega()
    ew = rel_l2(wg, wo)
    print(f"[periodic les omega] relL2={ew:.3e}", flush=True)
    assert ew <= 1e-12


def test_periodic_fp32_literal_vs_fp32_oracle(gpu):
    shape = (16, 16, 16)
    flags = np.ones(shape, np.int32)
    u = random_velocity_field(shape, 0.04, seed=3)
    g, o = _pair(flags, [], (None, *u), 30, collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY,
                 storage=abi.STORAGE_F32, arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE,
                 periodic=(1, 1, 1), tau=0.51, c_les=0.17)
    ef, er, eu = _report("periodic fp32 lit vs fp32 oracle", g, o, flags == 0)
    assert ef <= 1e-6 and eu <= 1e-5


def test_cavity_fast_fp64(gpu):
    """Reduced cumulant algebra in FP64 against the literal oracle."""
    flags, bc, ux = cavity_setup(10)
    g, o = _pair(flags, bc, (None, ux, None, None), 100, collision=abi.COLLISION_CUM2015,
                 storage=abi.STORAGE_F64, arith=abi.ARITH_FAST, compat=abi.COMPAT_REFERENCE, tau=C1_TAU)
    ef, er, eu = _report("cavity10 fp64 fast", g, o, flags == 0)
    assert ef <= 1e-10 and eu <= 1e-10


@pytest.mark.parametrize("storage,tol", [(abi.STORAGE_F32, 1e-4), (abi.STORAGE_F16DEV, 5e-3),
                                         (abi.STORAGE_BF16DEV, 5e-2), (abi.STORAGE_F16DEV_SR, 5e-3),
                                         (abi.STORAGE_BF16DEV_SR, 5e-2)],
                         ids=["f32", "f16dev", "bf16dev", "f16dev_sr", "bf16dev_sr"])
def test_periodic_reduced_precision_vs_fp64(gpu, storage, tol):
    shape = (16, 16, 16)
    flags = np.ones(shape, np.int32)
    u = random_velocity_field(shape, 0.04, seed=11)
    d = make_desc(flags, [], collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY, storage=storage,
                  arith=abi.ARITH_FAST, compat=abi.COMPAT_REFERENCE, periodic=(1, 1, 1), tau=0.51, c_les=0.17)
    dref = make_desc(flags, [], collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY,
                     storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE,
                     periodic=(1, 1, 1), tau=0.51, c_les=0.17)
    g, o = Lattice(d), OracleLattice(dref)
    g.init_equilibrium(None, *u)
    o.init_equilibrium(None, *u)
    g.step(100)
    o.step(100)
    mg, mo = g.macroscopic(), o.macroscopic()
    eu = rel_l2(np.stack(mg[1:]), np.stack(mo[1:]))
    er = rel_l2(mg[0] - 1.0, mo[0] - 1.0)
    print(f"[{abi.STORAGE_NAMES[storage]} vs fp64 oracle, 100 steps] relL2 u={eu:.3e} (rho-1)={er:.3e}", flush=True)
    assert eu <= tol


def test_conservation_periodic(gpu):
    shape = (16, 16, 16)
    flags = np.ones(shape, np.int32)
    u = random_velocity_field(shape, 0.04, seed=5)
    L = Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY,
                          storage=abi.STORAGE_F64, arith=abi.ARITH_FAST, compat=0, periodic=(1, 1, 1),
                          tau=0.51, c_les=0.17))
    L.init_equilibrium(None, *u)
    L.step(1)
    m0 = L.conservation()
    L.step(50)
    m1 = L.conservation()
    print(f"[conservation] start {m0} end {m1}", flush=True)
    assert abs(m1[0] - m0[0]) <= 1e-12 * abs(m0[0])
    assert np.all(np.abs(m1[1:] - m0[1:]) <= 1e-12 * abs(m0[0]))


@pytest.mark.parametrize("storage", [abi.STORAGE_F64, abi.STORAGE_F32, abi.STORAGE_F16DEV],
                         ids=["f64", "f32", "f16dev"])
def test_conservation_equals_host_sum(gpu, storage):
    """The device's FP64 mass/momentum reduction equals a brute-force host
    sum over the cells."""
    flags = np.ones((12, 10, 14), np.int32)
    L = Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY,
                          storage=storage, arith=abi.ARITH_FAST, compat=abi.COMPAT_REFERENCE, periodic=(1, 1, 1),
                          tau=0.55, c_les=0.17))
    u = random_velocity_field(flags.shape, 0.03, seed=17)
    L.init_equilibrium(1.0 + 0.1 * u[2], u[0], u[1], u[2])
    L.step(7)
    m = L.conservation()
    f = L.populations()
    fluid = flags != 0
    mass = f.sum()
    rho, vx, vy, vz = L.macroscopic()
    mom = np.array([(rho * v)[fluid].sum() for v in (vx, vy, vz)])
    scale = np.abs(rho[fluid]).sum()
    print(f"[conservation vs host, storage={storage}] mass {m[0]:.15g} host {mass:.15g}; "
          f"momentum {m[1:]} host {mom}", flush=True)
    assert abs(m[0] - mass) <= 1e-12 * mass
    # momentum: the host forms rho * u from the stored macroscopic fields, which
    # the kernel rounds in the storage's arithmetic (FP32 below FP64)
    tol = 1e-13 if storage == abi.STORAGE_F64 else 1e-6
    assert np.all(np.abs(m[1:] - mom) <= tol * scale)


def test_set_get_populations_roundtrip(gpu):
    flags, bc, ux = cavity_setup(6)
    L = Lattice(make_desc(flags, bc, collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F64,
                          arith=abi.ARITH_LITERAL, tau=C1_TAU))
    L.init_equilibrium(None, ux, None, None)
    rng = np.random.default_rng(0)
    for steps in (0, 1):
        L.step(steps)
        f = rng.uniform(0.0, 0.1, (27, *flags.shape))
        f[:, flags == 0] = 0.0
        L.set_populations(f)
        assert np.array_equal(L.populations(), f)


def test_set_populations_then_step_matches_oracle(gpu):
    flags, bc, ux = cavity_setup(6)
    kw = dict(collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL,
              compat=abi.COMPAT_REFERENCE, tau=C1_TAU)
    g, o = Lattice(make_desc(flags, bc, **kw)), OracleLattice(make_desc(flags, bc, **kw))
    g.init_equilibrium(None, ux, None, None)
    o.init_equilibrium(None, ux, None, None)
    g.step(3)
    o.step(3)
    rng = np.random.default_rng(1)
    f = o.populations() * (1 + 0.01 * rng.standard_normal((27, *flags.shape)))
    g.set_populations(f)
    o.set_populations(f)
    g.step(5)
    o.step(5)
    ef, er, eu = _report("set_populations + 5 steps", g, o, flags == 0)
    assert ef <= 1e-12


@pytest.mark.timeout(900)
def test_c1_cavity_64_2000_steps(gpu):
    """Config C1: 64^3 lid-driven cavity, Re = 400, 2000 steps, FP64,
    cumulant, reference compat tables: rel-L2(rho, u) <= 1e-12."""
    flags, bc, ux = cavity_setup(64)
    kw = dict(collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL,
              compat=abi.COMPAT_REFERENCE, tau=C1_TAU)
    g = Lattice(make_desc(flags, bc, **kw))
    g.init_equilibrium(None, ux, None, None)
    g.step(2000)
    g.sync()
    print(f"[c1] gpu done, {g.last_kernel_ms():.4f} ms/step", flush=True)
    o = OracleLattice(make_desc(flags, bc, **kw))
    o.init_equilibrium(None, ux, None, None)
    for k in range(10):
        o.step(200)
        print(f"[c1] oracle step {200 * (k + 1)}", flush=True)
    mg, mo = g.macroscopic(), o.macroscopic()
    fluid = flags != 0
    er = rel_l2(mg[0][fluid], mo[0][fluid])
    eu = rel_l2(np.stack([m[fluid] for m in mg[1:]]), np.stack([m[fluid] for m in mo[1:]]))
    print(f"[c1] relL2 rho={er:.3e} u={eu:.3e} max|u|={np.abs(mo[1]).max():.4f}", flush=True)
    assert er <= 1e-12 and eu <= 1e-12
