"""Outlet variants and the outlet omega damping of the HIP core vs the CPU
oracle, through the C ABI: OutletDensityEquilibrium, OutletNeighbourEquilibrium,
...Omega (the same rule plus omega damping) and ...OmegaRF (the backflow density
correction), on the channel with equilibrium inlet, ghost layers, walls and an
obstacle.  FP64 literal must match the oracle bit for bit (bar <= 1e-12 relative
L2, omega fields equal); fast arithmetic and reduced storage within the bounds
of the interpolated outlet's tests.
"""
import numpy as np
import pytest

from helpers import OracleLattice, bc_nodes, rel_l2
from lb_amd import Lattice, LbmError, abi, make_desc
from oracle import mesher_oracle as mo

pytestmark = pytest.mark.gpu

NU = 0.01
TAU = 3 * NU + 0.5
KINDS = [abi.BC_OUTLET_EQUILIBRIUM, abi.BC_OUTLET_NEIGHBOUR_EQ, abi.BC_OUTLET_NEIGHBOUR_EQ_RF]
IDS = ["density_eq", "neighbour_eq", "neighbour_eq_rf"]


def _run_pair(kind, steps, u0=0.04, oracle_storage=abi.STORAGE_F64, **kw):
    flags, bcd = mo.channel(24, 12, 10, outlet_kind=kind)
    bc = bc_nodes(bcd)
    print(f"[outlet pair] kind={kind} steps={steps} u0={u0} kw={kw}", flush=True)
    g = Lattice(make_desc(flags, bc, **kw))
    o = OracleLattice(make_desc(flags, bc, **{**kw, "storage": oracle_storage}))
    rng = np.random.default_rng(17)
    solid = flags == 0
    init = (np.where(solid, 1.0, 1.0 + 0.002 * rng.standard_normal(flags.shape)),
            np.where(solid, 0.0, u0 + 0.005 * rng.standard_normal(flags.shape)),
            np.where(solid, 0.0, 0.005 * rng.standard_normal(flags.shape)), None)
    g.init_equilibrium(*init)
    o.init_equilibrium(*init)
    g.step(steps)
    o.step(steps)
    g.sync()
    return flags, g, o


def _errors(tag, g, o, flags):
    fluid = flags != 0
    fg, fo = g.populations(), o.populations()
    mg, mo_ = g.macroscopic(), o.macroscopic()
    ef = rel_l2(fg[:, fluid], fo[:, fluid])
    er = rel_l2(mg[0][fluid], mo_[0][fluid])
    eu = rel_l2(np.stack([m[fluid] for m in mg[1:]]), np.stack([m[fluid] for m in mo_[1:]]))
    print(f"[{tag}] relL2 f={ef:.3e} rho={er:.3e} u={eu:.3e} max|df|={np.abs(fg - fo).max():.3e}", flush=True)
    return ef, er, eu


@pytest.mark.parametrize("kind", KINDS, ids=IDS)
@pytest.mark.parametrize("sponge", [None, (0.0, 0.05)], ids=["plain", "sponge"])
def test_channel_fp64_literal(gpu, kind, sponge):
    flags, g, o = _run_pair(kind, 120, collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY,
                            c_les=0.17, storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL,
                            compat=abi.COMPAT_REFERENCE, tau=TAU, nu=NU, sponge=sponge)
    ef, er, eu = _errors(f"outlet kind={kind} sponge={sponge}", g, o, flags)
    assert ef <= 1e-12 and er <= 1e-12 and eu <= 1e-12
    fluid = flags != 0
    assert np.array_equal(g.omega()[fluid], o.omega()[fluid])


def test_rf_backflow_branch_fp64_literal(gpu):
    """An initial backflow (u_x < 0 at the outlet) makes the RF density
    correction act; the Combined operator with its quirks runs beside it."""
    flags, g, o = _run_pair(abi.BC_OUTLET_NEIGHBOUR_EQ_RF, 40, u0=-0.03, collision=abi.COLLISION_COMBINED,
                            storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE,
                            tau=TAU, nu=NU, sponge=(0.0, 0.05))
    ef, er, eu = _errors("outlet rf backflow", g, o, flags)
    assert ef <= 1e-12 and er <= 1e-12 and eu <= 1e-12


def test_channel_fp32_literal_vs_fp32_oracle(gpu):
    flags, g, o = _run_pair(abi.BC_OUTLET_NEIGHBOUR_EQ, 80, oracle_storage=abi.STORAGE_F32,
                            collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY, c_les=0.17,
                            storage=abi.STORAGE_F32, arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE,
                            tau=TAU, nu=NU, sponge=(0.0, 0.05))
    fg, fo = g.populations(), o.populations()
    print(f"[outlet fp32 lit] max|df|={np.abs(fg - fo).max():.3e}", flush=True)
    assert np.array_equal(fg, fo)


@pytest.mark.parametrize("kind", KINDS, ids=IDS)
@pytest.mark.parametrize("storage,tol", [(abi.STORAGE_F64, 1e-8), (abi.STORAGE_F32, 2e-5),
                                         (abi.STORAGE_F16DEV_SR, 2e-3), (abi.STORAGE_BF16DEV, 1e-2)],
                         ids=["f64", "f32", "f16dev_sr", "bf16dev"])
def test_channel_fast_vs_fp64_oracle(gpu, kind, storage, tol):
    flags, g, o = _run_pair(kind, 100, collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY,
                            c_les=0.17, storage=storage, arith=abi.ARITH_FAST, compat=abi.COMPAT_REFERENCE,
                            tau=TAU, nu=NU, sponge=(0.0, 0.05))
    ef, er, eu = _errors(f"outlet fast kind={kind} storage={storage}", g, o, flags)
    assert eu <= tol and er <= tol


def test_sponge_needs_an_outlet_and_an_axis_normal(gpu):
    flags = np.ones((8, 8, 8), np.int32)
    with pytest.raises(LbmError):
        Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F32,
                          arith=abi.ARITH_FAST, periodic=(1, 1, 1), tau=0.6, sponge=(0.0, 0.05)))


@pytest.mark.parametrize("storage", [abi.STORAGE_F64, abi.STORAGE_F32], ids=["f64", "f32"])
def test_zou_he_inlet_literal_equals_oracle(gpu, storage):
    flags, bcd = mo.channel(24, 12, 10, inlet_kind=abi.BC_INLET_ZOU_HE)
    bc = bc_nodes(bcd)
    kw = dict(collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY, c_les=0.17, storage=storage,
              arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE, tau=TAU, nu=NU)
    g = Lattice(make_desc(flags, bc, **kw))
    o = OracleLattice(make_desc(flags, bc, **kw))
    solid = flags == 0
    rng = np.random.default_rng(23)
    init = (None, np.where(solid, 0.0, 0.03 + 0.004 * rng.standard_normal(flags.shape)), None, None)
    g.init_equilibrium(*init)
    o.init_equilibrium(*init)
    g.step(100)
    o.step(100)
    fg, fo = g.populations(), o.populations()
    print(f"[zou-he storage={storage}] max|df|={np.abs(fg - fo).max():.3e}", flush=True)
    assert np.array_equal(fg, fo)


def test_zou_he_inlet_fast_vs_fp64_oracle(gpu):
    flags, bcd = mo.channel(24, 12, 10, inlet_kind=abi.BC_INLET_ZOU_HE)
    bc = bc_nodes(bcd)
    kw = dict(collision=abi.COLLISION_COMBINED, closure=abi.CLOSURE_SMAGORINSKY, c_les=0.17,
              compat=abi.COMPAT_REFERENCE, tau=TAU, nu=NU)
    g = Lattice(make_desc(flags, bc, storage=abi.STORAGE_F32, arith=abi.ARITH_FAST, **kw))
    o = OracleLattice(make_desc(flags, bc, storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, **kw))
    init = (None, np.where(flags == 0, 0.0, 0.03), None, None)
    g.init_equilibrium(*init)
    o.init_equilibrium(*init)
    g.step(100)
    o.step(100)
    ef, er, eu = _errors("zou-he fast f32 combined", g, o, flags)
    assert eu <= 2e-5 and er <= 2e-5
