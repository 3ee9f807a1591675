"""ErrorQuadratic on the device (lbm_error_quadratic) against the oracle's
restatement of ErrorQuadratic.h:33-90, and its NaN report.

The oracle sums the per-cell terms sequentially in double; the device sums
them per block, then across blocks: the norms agree to FP64 summation
rounding (1e-12 relative), and the per-cell terms are formed the same way
in both (the arithmetic type, no contraction).
"""
import numpy as np
import pytest

from helpers import OracleLattice, bc_nodes, cavity_setup, random_velocity_field
from lb_amd import Lattice, LbmError, abi, make_desc
from oracle import mesher_oracle as mo

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", [abi.STORAGE_F64, abi.STORAGE_F32], ids=["f64", "f32"])
def test_error_norm_equals_oracle_over_calls(gpu, storage):
    flags, bcd = mo.channel(24, 12, 10)
    bc = bc_nodes(bcd)
    kw = dict(collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY, c_les=0.17, storage=storage,
              arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE, tau=0.53, nu=0.01)
    g = Lattice(make_desc(flags, bc, **kw))
    o = OracleLattice(make_desc(flags, bc, **kw))
    rng = np.random.default_rng(41)
    solid = flags == 0
    ux = np.where(solid, 0.0, 0.04 + 0.004 * rng.standard_normal(flags.shape))
    uy = np.where(solid, 0.0, 0.004 * rng.standard_normal(flags.shape))
    for L in (g, o):
        L.init_equilibrium(None, ux, uy, None)
    for k in range(4):
        eg, eo = g.error_quadratic(), o.error_quadratic()
        print(f"[error {abi.STORAGE_NAMES[storage]}] call {k}: gpu {eg:.15e} oracle {eo:.15e}", flush=True)
        assert abs(eg - eo) <= 1e-12 * abs(eo)
        g.step(7)
        o.step(7)
    assert (flags == 1).sum() < (flags != 0).sum()   # the norm skips cells with flag != 1


def test_error_norm_restarts_with_initialisation(gpu):
    flags, bc, ux = cavity_setup(10)
    g = Lattice(make_desc(flags, bc, storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, tau=0.6))
    g.init_equilibrium(None, ux, None, None)
    g.step(20)
    e0 = g.error_quadratic()
    g.step(5)
    assert g.error_quadratic() != e0
    g.init_equilibrium(None, ux, None, None)
    g.step(20)
    assert g.error_quadratic() == e0     # u_prev is zero again, the state is the same


def test_error_norm_reports_nan(gpu):
    """A NaN in the state makes the sums NaN: LBM_E_NAN, and lbm_sync keeps
    reporting it (the reference loop breaks on a NaN norm)."""
    n = 12
    flags = np.ones((n, n, n), np.int32)
    g = Lattice(make_desc(flags, [], storage=abi.STORAGE_F32, periodic=(1, 1, 1), tau=0.6))
    u = random_velocity_field(flags.shape, 0.02, seed=3)
    u[0][3, 4, 5] = np.nan
    g.init_equilibrium(None, *u)
    g.step(2)
    with pytest.raises(LbmError) as ei:
        g.error_quadratic()
    assert "LBM_E_NAN" in str(ei.value)
    with pytest.raises(LbmError):
        g.sync()


@pytest.mark.parametrize("compat", [0, abi.COMPAT_INLETRHO], ids=["corrected", "reference"])
@pytest.mark.parametrize("side", [-1, 1], ids=["minus_x", "plus_x"])
def test_equilibrium_inlet_both_modes_and_sides_equal_oracle(gpu, compat, side):
    """The inlet-density extrapolation (InletVelocityEquilibrium.h:80-95) in
    the corrected form and the reference's (LBM_COMPAT_INLETRHO), on inlets
    facing -x and +x, FP64 literal bit for bit against the oracle."""
    nx, n = 16, 6
    xb = 0 if side < 0 else nx - 1
    flags = np.ones((n, n, nx), np.int32)
    flags[:, :, xb] = -1
    bc = bc_nodes([dict(kind=abi.BC_INLET_EQUILIBRIUM, x=xb, y=y, z=z, normal=(float(side), 0.0, 0.0),
                        velocity=(-0.03 * side, 0.0, 0.0), regular=1) for z in range(n) for y in range(n)])
    kw = dict(collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY, c_les=0.17,
              storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=compat, periodic=(0, 1, 1),
              tau=0.56, nu=0.02)
    g = Lattice(make_desc(flags, bc, **kw))
    o = OracleLattice(make_desc(flags, bc, **kw))
    x = np.arange(nx)
    ux = np.broadcast_to(-0.03 * side + 0.004 * np.sin(0.5 * x) + 0.001 * x * x / nx, flags.shape)
    rho = np.broadcast_to(1 + 0.01 * np.cos(0.3 * x), flags.shape)
    for L in (g, o):
        L.init_equilibrium(rho, ux, None, None)
        L.step(25)
    fg, fo = g.populations(), o.populations()
    print(f"[inlet compat={compat} side={side}] max|df| {np.abs(fg - fo).max():.3e}", flush=True)
    assert np.array_equal(fg, fo)
