"""Wall force (drag/lift) of lbm_step_diagnostics: against the oracle's
post-collision populations summed on the host, and against the momentum
balance of the fluid (true opposite table, every solid-adjacent fluid cell a
wall node: the fluid's momentum changes by exactly minus the wall force)."""
import numpy as np
import pytest

from helpers import OracleLattice, bc_nodes, tables
from lb_amd import Lattice, abi, make_desc

pytestmark = pytest.mark.gpu


def _obstacle_box(nx=24, ny=16, nz=16):
    flags = np.ones((nz, ny, nx), np.int32)
    flags[5:10, 6:11, 8:12] = 0          # box obstacle
    flags[7:9, 10:12, 12:14] = 0         # a bump on it
    solid = flags == 0
    near = np.zeros_like(solid)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                near |= np.roll(solid, (dz, dy, dx), axis=(0, 1, 2))
    walls = np.argwhere(near & ~solid)
    flags[near & ~solid] = 5
    bc = bc_nodes([dict(kind=abi.BC_WALL, x=int(x), y=int(y), z=int(z)) for z, y, x in walls])
    return flags, bc, walls


def _host_force(fpost, flags, walls, compat):
    c, crev, _ = tables(compat)
    nz, ny, nx = flags.shape
    F = np.zeros(3)
    for z, y, x in walls:
        for q in range(1, 27):
            sx, sy, sz = (x - c[q, 0]) % nx, (y - c[q, 1]) % ny, (z - c[q, 2]) % nz
            if flags[sz, sy, sx] == 0:
                opp = int(np.where((c == -c[q]).all(1))[0][0])
                F -= c[q] * (fpost[crev[q], z, y, x] + fpost[opp, z, y, x])
    return F


@pytest.mark.parametrize("compat", [abi.COMPAT_REFERENCE, 0], ids=["reference", "corrected"])
def test_wall_force_matches_oracle_post_collision_sum(gpu, compat):
    flags, bc, walls = _obstacle_box()
    kw = dict(collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=compat,
              periodic=(1, 1, 1), tau=0.56)
    g = Lattice(make_desc(flags, bc, **kw))
    o = OracleLattice(make_desc(flags, bc, **kw))
    ux = np.where(flags == 0, 0.0, 0.04)
    for L in (g, o):
        L.init_equilibrium(None, ux, None, None)
    g.step(20)
    o.step(20)
    _, _, F = g.step_diagnostics()
    o.step(1)
    Fh = _host_force(o.post_populations(), flags, walls, compat)
    print(f"[drag compat={compat}] gpu {F} host {Fh}", flush=True)
    assert F[0] > 0
    # the GPU sums per block, the host sequentially: components that cancel
    # between links carry the rounding of the O(1e-2) link terms
    assert np.allclose(F, Fh, rtol=1e-10, atol=1e-12)


def test_fluid_momentum_balances_wall_force(gpu):
    flags, bc, _ = _obstacle_box()
    g = Lattice(make_desc(flags, bc, collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F64, arith=abi.ARITH_FAST,
                          compat=0, periodic=(1, 1, 1), tau=0.56))
    g.init_equilibrium(None, np.where(flags == 0, 0.0, 0.04), None, None)
    g.step(5)
    m1, J1, F1 = g.step_diagnostics()
    m2, J2, F2 = g.step_diagnostics()
    print(f"[momentum balance] dJ {J2 - J1} F {F1} dm {m2 - m1:.3e}", flush=True)
    assert abs(m2 - m1) <= 1e-11 * m1
    assert np.allclose(J2 - J1, -F1, rtol=1e-9, atol=1e-13)
