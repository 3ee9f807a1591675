"""Pins of the CPU oracle against what the mathematics fixes (no GPU).

Each test is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign or index, a transposed operand) fails it: closed forms of the
D3Q27 lattice, conservation and fixed points of the collision, a brute-force
definition of the central moments, the viscosity measured from shear-wave
decay, the equivalence of the reference's ghost-layer periodicity with index
wrap, and the product collision arithmetic compiled for the host.
"""
import ctypes
import itertools

import numpy as np
import pytest

from helpers import (OracleLattice, cavity_setup, host_collide_lib, oracle_lib, random_velocity_field, rel_l2,
                     tables)
from lb_amd import abi, make_desc

dp = ctypes.POINTER(ctypes.c_double)


def feq(rho, u, compat=0):
    out = np.zeros(27)
    oracle_lib().orc_cell_feq(compat, 1, rho, *u, abi.ptr(out))
    return out


# --- lattice tables -------------------------------------------------------

def test_velocity_set_and_weights():
    c, crev, w = tables(0)
    assert len({tuple(v) for v in c}) == 27 and c.min() == -1 and c.max() == 1
    assert np.isclose(w.sum(), 1.0, atol=1e-15)
    assert np.allclose(w @ c, 0.0, atol=1e-16)
    assert np.allclose(np.einsum("q,qa,qb->ab", w, c, c), np.eye(3) / 3, atol=1e-16)
    # fourth-order isotropy: sum w c_a c_b c_c c_d = (d_ab d_cd + d_ac d_bd + d_ad d_bc) / 9
    d = np.eye(3)
    iso = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d)) / 9
    assert np.allclose(np.einsum("q,qa,qb,qc,qd->abcd", w, c, c, c, c), iso, atol=1e-16)
    # corrected reverse table is the true opposite
    assert np.array_equal(c[crev], -c)


def test_reference_quirks_q1_q2():
    c, crev_ref, w_ref = tables(abi.COMPAT_REFERENCE)
    _, crev, w = tables(0)
    diff = np.nonzero(crev_ref != crev)[0]
    assert list(diff) == [20] and crev_ref[20] == 21 and crev[20] == 19
    # float-rounded weights: equal to the float32 quotient, not the exact value
    assert w_ref[0] == float(np.float32(8) / np.float32(27)) and w_ref[0] != 8.0 / 27.0
    assert abs(w_ref.sum() - 1.0) < 1e-7 and w_ref.sum() != 1.0


def test_equilibrium_moments_closed_form():
    c, _, w = tables(0)
    rho, u = 1.03, np.array([0.04, -0.03, 0.02])
    f = feq(rho, u)
    assert np.isclose(f.sum(), rho, rtol=1e-15)
    assert np.allclose(f @ c, rho * u, atol=1e-16)
    # second-order polynomial equilibrium of D3Q27: sum c c f_eq = rho (u u + I/3)
    assert np.allclose(np.einsum("q,qa,qb->ab", f, c, c), rho * (np.outer(u, u) + np.eye(3) / 3), atol=1e-15)


# --- collision ------------------------------------------------------------

@pytest.mark.parametrize("coll", [abi.COLLISION_SRT, abi.COLLISION_CUM2015])
@pytest.mark.parametrize("dbl", [1, 0])
def test_collision_conserves_mass_and_momentum(coll, dbl):
    c, _, _ = tables(0)
    rng = np.random.default_rng(3)
    O = oracle_lib()
    for _ in range(20):
        f = feq(1 + 0.02 * rng.standard_normal(), 0.04 * rng.standard_normal(3)) * (1 + 0.03 * rng.standard_normal(27))
        if not dbl:
            f = f.astype(np.float32).astype(np.float64)
        rho = f.sum()
        u = f @ c / rho
        out = np.zeros(27)
        O.orc_cell_collide(coll, 0, dbl, abi.ptr(f), rho, *u, 1.7, 1, abi.ptr(out))
        tol = 1e-14 if dbl else 2e-6
        assert abs(out.sum() - rho) < tol
        assert np.allclose(out @ c, rho * u, atol=tol)


@pytest.mark.parametrize("coll", [abi.COLLISION_SRT, abi.COLLISION_CUM2015])
def test_rest_equilibrium_is_fixed_point(coll):
    f = feq(1.0, (0.0, 0.0, 0.0))
    out = np.zeros(27)
    oracle_lib().orc_cell_collide(coll, 0, 1, abi.ptr(f), 1.0, 0.0, 0.0, 0.0, 1.9, 1, abi.ptr(out))
    assert np.allclose(out, f, atol=2e-16)


def test_collision_with_omega_one_relaxes_stress_to_equilibrium():
    """omega = 1 relaxes every relaxable moment to equilibrium: the deviatoric
    second central moments vanish after collision (both operators)."""
    c, _, _ = tables(0)
    rng = np.random.default_rng(4)
    f = feq(1.0, (0.02, -0.01, 0.03)) * (1 + 0.05 * rng.standard_normal(27))
    rho = f.sum()
    u = f @ c / rho
    for coll in (abi.COLLISION_SRT, abi.COLLISION_CUM2015):
        out = np.zeros(27)
        oracle_lib().orc_cell_collide(coll, 0, 1, abi.ptr(f), rho, *u, 1.0, 1, abi.ptr(out))
        cc = c - u
        k2 = np.einsum("q,qa,qb->ab", out, cc, cc)
        dev = k2 - np.trace(k2) / 3 * np.eye(3)
        assert np.abs(dev).max() < 1e-5, coll


def test_central_moment_transform_matches_definition():
    """The three-pass transform equals the brute-force definition
    k_ijl = sum_q (cx-ux)^i (cy-uy)^j (cz-uz)^l f_q, and its inverse undoes it."""
    c, _, _ = tables(0)
    rng = np.random.default_rng(5)
    O = oracle_lib()
    for _ in range(10):
        f = rng.uniform(0.0, 0.3, 27)
        u = 0.05 * rng.standard_normal(3)
        k = np.zeros(27)
        O.orc_cell_central_moments(abi.ptr(f), *u, abi.ptr(k))
        cc = c - u
        for i, j, l in itertools.product(range(3), repeat=3):
            ref = np.sum(cc[:, 0] ** i * cc[:, 1] ** j * cc[:, 2] ** l * f)
            assert abs(k[9 * i + 3 * j + l] - ref) < 1e-14, (i, j, l)
        back = np.zeros(27)
        O.orc_cell_from_central_moments(abi.ptr(k), *u, abi.ptr(back))
        assert np.allclose(back, f, atol=1e-15)


def test_cumulant_relaxes_shear_at_rate_omega():
    """Off-diagonal second cumulants relax as C* = (1 - omega) C (Eq. G2015 55-57)."""
    c, _, _ = tables(0)
    rng = np.random.default_rng(6)
    f = feq(1.0, (0.01, 0.02, -0.01)) * (1 + 0.02 * rng.standard_normal(27))
    rho = f.sum()
    u = f @ c / rho
    om = 1.3
    out = np.zeros(27)
    oracle_lib().orc_cell_collide(abi.COLLISION_CUM2015, 0, 1, abi.ptr(f), rho, *u, om, 1, abi.ptr(out))
    cc = c - u
    k_in = np.einsum("q,qa,qb->ab", f, cc, cc)
    k_out = np.einsum("q,qa,qb->ab", out, cc, cc)
    for a, b in ((0, 1), (0, 2), (1, 2)):
        assert abs(k_out[a, b] - (1 - om) * k_in[a, b]) < 1e-15


# --- Smagorinsky ----------------------------------------------------------

def test_smagorinsky_special_cases():
    O = oracle_lib()
    tau = 0.52
    f = feq(1.0, (0.02, 0.01, 0.0))
    # equilibrium: Pi = 0 -> omega = 1/tau
    w = O.orc_cell_omega_les(0, 1, abi.ptr(f), 1.0, 0.02, 0.01, 0.0, tau, 0.173)
    assert abs(w - 1 / tau) < 1e-14
    # C = 0 -> omega = 1/tau for any f
    g = f * (1 + 0.05 * np.random.default_rng(0).standard_normal(27))
    w = O.orc_cell_omega_les(0, 1, abi.ptr(g), 1.0, 0.02, 0.01, 0.0, tau, 0.0)
    assert abs(w - 1 / tau) < 1e-14


def test_smagorinsky_known_stress():
    """Perturb the four (+-1, +-1, 0) populations so that Pi = s (e_x e_y + e_y e_x);
    |Pi|_F = sqrt(2) s and tau_LES = tau/2 + sqrt(tau^2 + 18 C |Pi| / rho)/2."""
    c, _, _ = tables(0)
    O = oracle_lib()
    rho, tau, C, s = 1.0, 0.51, 0.173, 3e-4
    f = feq(rho, (0.0, 0.0, 0.0))
    for q in range(27):
        if c[q, 2] == 0 and c[q, 0] != 0 and c[q, 1] != 0:
            f[q] += 0.25 * s * c[q, 0] * c[q, 1]
    pi = np.einsum("q,qa,qb->ab", f - feq(rho, (0, 0, 0)), c, c)
    assert np.allclose(pi, s * np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]]), atol=1e-18)
    w = O.orc_cell_omega_les(0, 1, abi.ptr(f), rho, 0.0, 0.0, 0.0, tau, C)
    tau_les = tau / 2 + 0.5 * np.sqrt(tau * tau + 18 * C * np.sqrt(2) * s / rho)
    assert abs(w - 1 / tau_les) < 1e-13


# --- whole-step physics ---------------------------------------------------

@pytest.mark.parametrize("coll", [abi.COLLISION_SRT, abi.COLLISION_CUM2015])
def test_shear_wave_decay_gives_viscosity(coll):
    """u_y(x) = A sin(2 pi x / L) decays as exp(-nu k^2 t) with nu = (tau - 1/2)/3."""
    L, tau, steps = 32, 0.8, 300
    shape = (2, 2, L)
    x = np.arange(L)
    uy = np.broadcast_to(1e-3 * np.sin(2 * np.pi * x / L), shape).copy()
    d = make_desc(np.ones(shape, np.int32), [], collision=coll, storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL,
                  compat=0, periodic=(1, 1, 1), tau=tau)
    o = OracleLattice(d)
    o.init_equilibrium(None, None, uy, None)
    o.step(50)
    a0 = np.abs(o.macroscopic()[2][0, 0]).max()
    o.step(steps)
    a1 = np.abs(o.macroscopic()[2][0, 0]).max()
    k = 2 * np.pi / L
    nu_meas = -np.log(a1 / a0) / (k * k * steps)
    nu = (tau - 0.5) / 3
    assert abs(nu_meas / nu - 1) < 0.01, (nu_meas, nu)


def test_periodic_ghost_layer_equals_index_wrap():
    """Away from walls, the reference's ghost-layer Periodic policy
    (Periodic.h:43-138) reproduces index-wrap periodicity bit for bit: with
    ghost rows y = 0 (partner y = n) and y = n + 1 (partner y = 1), the interior
    of a y-periodic box equals the index-wrapped box.  (At walls the two
    differ: a regular ghost node, flag 10, gets no bounce-back there.)"""
    from helpers import bc_nodes
    from oracle import mesher_oracle as mo

    n = 8
    shape_w = (6, n, 7)
    rng = np.random.default_rng(9)
    flags_w = np.ones(shape_w, np.int32)
    g_g = np.ones((6, n + 2, 7), np.int32)
    flags_g, walls_g = mo.meshing_boundary_wall(g_g)
    sh = flags_g.shape
    m0 = mo.cuboid_mask(sh, (-50, -0.6, -50), (50, -0.4, 50), (-50, -0.6, 50), additional=1.0)
    m1 = mo.cuboid_mask(sh, (-50, n + 0.4, -50), (50, n + 0.6, 50), (-50, n + 0.4, 50), additional=1.0)
    per = mo.meshing_periodic(flags_g, m0, -3, (0.0, -1.0, 0.0), n)
    per += mo.meshing_periodic(flags_g, m1, -3, (0.0, 1.0, 0.0), 1)
    assert len(per) == 2 * 6 * 7 and not mo.compile_wall(flags_g, walls_g)
    bc_g = bc_nodes([dict(kind=mo.BC_PERIODIC_GHOST, **p) for p in per])
    ux = 0.02 * rng.standard_normal(shape_w)
    uxg = np.concatenate([ux[:, -1:, :], ux, ux[:, :1, :]], axis=1)
    kw = dict(collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL,
              compat=abi.COMPAT_REFERENCE, tau=0.6)
    ow = OracleLattice(make_desc(flags_w, [], periodic=(1, 1, 1), **kw))
    og = OracleLattice(make_desc(flags_g, bc_g, periodic=(1, 0, 1), **kw))
    ow.init_equilibrium(None, ux, None, None)
    og.init_equilibrium(None, uxg, None, None)
    ow.step(25)
    og.step(25)
    assert np.array_equal(ow.populations(), og.populations()[:, :, 1:-1, :])


def test_closed_box_mass_drift_is_the_crev_quirk():
    """In a closed box with half-way bounce-back on every wall, mass is
    conserved exactly with the true opposite table and drifts with the
    reference's c_rev[20] = 21 (quirk Q1)."""
    from helpers import bc_nodes
    from oracle import mesher_oracle as mo

    g = np.zeros((10, 10, 10), np.int32)
    g[1:-1, 1:-1, 1:-1] = 1
    flags, walls = mo.meshing_boundary_wall(g)
    bc = bc_nodes([dict(kind=mo.BC_WALL, x=a, y=b, z=c, regular=r) for a, b, c, r in walls])
    u = [a * (flags > 0) for a in random_velocity_field(flags.shape, 0.02, 1)]
    fluid = flags > 0

    def mass_change(compat):
        o = OracleLattice(make_desc(flags, bc, collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F64,
                                    arith=abi.ARITH_LITERAL, compat=compat, tau=0.524))
        o.init_equilibrium(None, *u)
        o.step(1)
        m0 = o.macroscopic()[0][fluid].sum()
        o.step(40)
        return o.macroscopic()[0][fluid].sum() - m0

    assert abs(mass_change(abi.COMPAT_REFERENCE)) > 1e-3
    assert abs(mass_change(abi.COMPAT_F32CONST)) < 1e-12
    assert abs(mass_change(0)) < 1e-12


def test_error_quadratic_known_values():
    """ErrorQuadratic with the y-z typo of ErrorQuadratic.h:43-44: for a field
    with uz = 0 the third term uses uy - uz_prev = uy."""
    shape = (4, 4, 4)
    flags = np.ones(shape, np.int32)
    d = make_desc(flags, [], collision=abi.COLLISION_SRT, storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL,
                  compat=0, periodic=(1, 1, 1), tau=0.8)
    o = OracleLattice(d)
    o.init_equilibrium(None, 0.01, 0.02, 0.0)
    e = o.error_quadratic()  # previous u = 0: err = sqrt(sum sqrt(ux^2 + 2 uy^2)) / sqrt(sum |u|)
    n = 64
    expect = np.sqrt(n * np.sqrt(0.01 ** 2 + 2 * 0.02 ** 2)) / np.sqrt(n * np.sqrt(0.01 ** 2 + 0.02 ** 2))
    assert abs(e - expect) < 1e-12
    # u_prev now equals u, yet the typo keeps the third term (uy - uz_prev)^2 = uy^2 alive
    e2 = o.error_quadratic()
    assert abs(e2 - np.sqrt(0.02) / np.sqrt(np.sqrt(0.01 ** 2 + 0.02 ** 2))) < 1e-12


# --- product arithmetic on the host --------------------------------------

def test_product_collision_arithmetic_is_bitwise_the_oracle():
    """lb_amd/csrc/collide.hpp (literal variant) compiled for the host gives
    the oracle's numbers bit for bit; the reduced variant agrees to ~1 ulp."""
    H, O = host_collide_lib(), oracle_lib()
    c, _, _ = tables(0)
    rng = np.random.default_rng(11)
    for _ in range(50):
        f0 = feq(1 + 0.02 * rng.standard_normal(), 0.05 * rng.standard_normal(3)) * (1 + 0.02 * rng.standard_normal(27))
        for dbl in (1, 0):
            f = f0 if dbl else f0.astype(np.float32).astype(np.float64)
            m = np.zeros(4)
            H.host_moments(dbl, abi.ptr(f), abi.ptr(m))
            for coll in (0, 1):
                for compat in (0, abi.COMPAT_REFERENCE):
                    a, b, fast = np.zeros(27), np.zeros(27), np.zeros(27)
                    O.orc_cell_collide(coll, compat, dbl, abi.ptr(f), *m, 1.9, 1, abi.ptr(a))
                    H.host_collide(coll, 0, dbl, compat, abi.ptr(f), *m, 1.9, abi.ptr(b))
                    H.host_collide(coll, 1, dbl, compat, abi.ptr(f), *m, 1.9, abi.ptr(fast))
                    assert np.array_equal(a, b)
                    assert np.abs(fast - a).max() < (1e-15 if dbl else 3e-7)
            wo = O.orc_cell_omega_les(abi.COMPAT_REFERENCE, dbl, abi.ptr(f), *m, 0.5015, 0.173)
            wh = H.host_omega_les(0, dbl, abi.COMPAT_REFERENCE, abi.ptr(f), *m, 0.5015, 0.173)
            assert wo == wh


# --- inlet / outlet / ghost-layer rules -----------------------------------

def _channel(nx, ny, nz, **kw):
    from helpers import bc_nodes
    from oracle import mesher_oracle as mo

    flags, bc = mo.channel(nx, ny, nz, **kw)
    return flags, bc, bc_nodes(bc)


def test_ghost_layer_with_walls_equals_index_wrap_for_z_invariant_flow():
    """Ghost layers beside walls: irregular ghost nodes bounce their own
    post-collision populations back off solids (Periodic.h:92-133).  For a
    z-invariant flow in a y-walled channel this must equal index wrap along z
    exactly (true opposite table: the ghost rule DF(crev v) = DP(v) and the
    wall rule DF(q) = DP(crev q) coincide only when crev is an involution)."""
    from helpers import bc_nodes
    from oracle import mesher_oracle as mo

    nx, ny, nzi = 9, 8, 4
    g = np.ones((nzi + 2, ny, nx), np.int32)
    g[:, 0, :] = 0
    g[:, ny - 1, :] = 0
    flags_g, walls_g = mo.meshing_boundary_wall(g)
    sh = flags_g.shape
    m0 = mo.cuboid_mask(sh, (-50, -50, 0.2), (50, 50, 0.8), (-50, 50, 0.8), additional=0.0)
    m1 = mo.cuboid_mask(sh, (-50, -50, nzi + 1.2), (50, 50, nzi + 1.8), (-50, 50, nzi + 1.8), additional=0.0)
    per = mo.meshing_periodic(flags_g, m0, -3, (0.0, 0.0, -1.0), nzi)
    per += mo.meshing_periodic(flags_g, m1, -3, (0.0, 0.0, 1.0), 1)
    assert any(p["regular"] == 0 and p["y"] == 1 for p in per)
    walls_g = mo.compile_wall(flags_g, walls_g)
    bc_g = bc_nodes([dict(kind=mo.BC_WALL, x=w[0], y=w[1], z=w[2], regular=w[3]) for w in walls_g] +
                    [dict(kind=mo.BC_PERIODIC_GHOST, **p) for p in per])
    gw = g[1:-1]
    flags_w, walls_w = mo.meshing_boundary_wall(gw)
    bc_w = bc_nodes([dict(kind=mo.BC_WALL, x=w[0], y=w[1], z=w[2], regular=w[3]) for w in walls_w])
    y = np.arange(ny)[None, :, None]
    x = np.arange(nx)[None, None, :]
    ux = np.broadcast_to(0.03 * np.sin(np.pi * y / ny) + 0.01 * np.cos(2 * np.pi * x / nx), gw.shape)
    uy = np.broadcast_to(0.01 * np.sin(2 * np.pi * x / nx) * (y > 0) * (y < ny - 1), gw.shape)
    kw = dict(collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY, c_les=0.17,
              storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=0, tau=0.56)
    ow = OracleLattice(make_desc(flags_w, bc_w, periodic=(1, 0, 1), **kw))
    og = OracleLattice(make_desc(flags_g, bc_g, periodic=(1, 0, 0), **kw))
    solid_w = flags_w == 0
    ow.init_equilibrium(None, np.where(solid_w, 0, ux), np.where(solid_w, 0, uy), None)
    pad = lambda a: np.concatenate([a[-1:], a, a[:1]], axis=0)  # noqa: E731
    og.init_equilibrium(None, pad(np.where(solid_w, 0, ux)), pad(np.where(solid_w, 0, uy)), None)
    ow.step(30)
    og.step(30)
    fw, fg = ow.populations(), og.populations()
    assert np.abs(fw).max() > 0 and np.array_equal(fw, fg[:, 1:-1])
    assert np.array_equal(fg[:, 0], fw[:, -1]) and np.array_equal(fg[:, -1], fw[:, 0])


def test_uniform_flow_through_inlet_outlet_channel_is_steady():
    """A uniform equilibrium flow u0 at rho = 1 is a fixed point of the
    equilibrium inlet (rho_b extrapolates to 1, state = f_eq(u0, 1)), the
    interpolated outlet (interpolating equal values, density reset to 1) and
    the ghost layers (copying equal values)."""
    u0 = 0.04
    flags, _, bc = _channel(12, 6, 6, u_in=u0, obstacle=False, walls=False)
    assert (flags == -1).sum() and (flags == -2).sum() and (flags == -3).sum()
    o = OracleLattice(make_desc(flags, bc, periodic=(0, 1, 0), collision=abi.COLLISION_CUM2015,
                                storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=0, tau=0.6, nu=0.1 / 3))
    u = np.float64(np.float32(u0))
    o.init_equilibrium(None, np.full(flags.shape, u), None, None)
    o.step(20)
    f = o.populations()
    expect = feq(1.0, (u, 0.0, 0.0))
    assert np.abs(f - expect[:, None, None, None]).max() < 1e-15


def test_outlet_resets_density_and_inlet_holds_equilibrium():
    """After every step the outlet nodes hold total mass 1 exactly (exact
    weights: f - f_eq(u, rho) + f_eq(u, 1) with sum f_eq(u, r) = r,
    OutletDensityInterpolated.h:244-246), and every inlet node holds
    f_eq(u_b, rho_b) with rho_b = rho(x+1) - 3 nu (u_b - 2 u(x+1) + u(x+2))
    from the neighbours' moments entering the step (InletVelocityEquilibrium.h:80-92
    with the corrected coefficient, written here as a second difference)."""
    nu = 0.02
    flags, bcd, bc = _channel(16, 8, 6, u_in=0.05)
    o = OracleLattice(make_desc(flags, bc, collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY,
                                c_les=0.17, storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL,
                                compat=0, tau=3 * nu + 0.5, nu=nu))
    rng = np.random.default_rng(3)
    solid = flags == 0
    o.init_equilibrium(1 + 0.01 * rng.standard_normal(flags.shape), np.where(solid, 0, 0.03), None, None)
    outlets = [d for d in bcd if d["kind"] == 3 and flags[d["z"], d["y"], d["x"]] == -2 and 0 < d["z"] < 5]
    inlets = [d for d in bcd if d["kind"] == 2]
    assert outlets and inlets
    for _ in range(4):
        rho0, ux0, _, _ = o.macroscopic()
        o.step(1)
        f = o.populations()
        for d in outlets:
            assert abs(f[:, d["z"], d["y"], d["x"]].sum() - 1.0) < 1e-14
        for d in inlets:
            x, y, z = d["x"], d["y"], d["z"]
            ub = np.float64(np.float32(d["velocity"][0]))
            rb = rho0[z, y, x + 1] - 3 * nu * (ub - 2 * ux0[z, y, x + 1] + ux0[z, y, x + 2])
            got = f[:, z, y, x]
            assert abs(got.sum() - rb) < 1e-13
            assert np.abs(got - feq(rb, (ub, 0.0, 0.0))).max() < 1e-15


# --- time averaging -------------------------------------------------------

def test_moment_average_is_the_mean_of_the_snapshots():
    """MomentTimeAvg (MomentTimeAvg.h:19-77): the average equals the brute-force
    mean of the rho, u fields read after every counted step, summed in the same
    order; the count is the number of adds."""
    from oracle import mesher_oracle as mo
    from helpers import bc_nodes
    flags, bcd = mo.channel(16, 8, 6)
    bc = bc_nodes(bcd)
    o = OracleLattice(make_desc(flags, bc, collision=abi.COLLISION_COMBINED, storage=abi.STORAGE_F64,
                                arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE, tau=0.53, nu=0.01))
    rng = np.random.default_rng(4)
    o.init_equilibrium(None, np.where(flags == 0, 0.0, 0.04 + 0.003 * rng.standard_normal(flags.shape)), None, None)
    o.step(3)
    acc = [np.zeros(flags.shape) for _ in range(4)]
    for k in range(6):
        o.step(1)
        o.moment_add()
        for a, m in zip(acc, o.macroscopic()):
            a += m
    avg, n = o.moment_avg()
    assert n == 6
    for a, b in zip(avg, acc):
        assert np.array_equal(a, b / 6)
    # a uniform steady flow averages to itself
    flags2 = np.ones((4, 4, 4), np.int32)
    o2 = OracleLattice(make_desc(flags2, [], collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F64,
                                 arith=abi.ARITH_LITERAL, compat=0, periodic=(1, 1, 1), tau=0.6))
    o2.init_equilibrium(None, 0.03, -0.01, 0.02)
    o2.step_moment_add(5)
    avg2, n2 = o2.moment_avg()
    assert n2 == 5
    assert np.allclose(avg2[0], 1.0, atol=1e-14) and np.allclose(avg2[1], 0.03, atol=1e-15)
    assert np.allclose(avg2[2], -0.01, atol=1e-15) and np.allclose(avg2[3], 0.02, atol=1e-15)


# --- outlet variants and omega damping ------------------------------------

@pytest.mark.parametrize("kind", [abi.BC_OUTLET_NEIGHBOUR_EQ, abi.BC_OUTLET_NEIGHBOUR_EQ_RF],
                         ids=["neighbour_eq", "neighbour_eq_rf"])
def test_uniform_flow_is_a_fixed_point_of_the_neighbour_outlets(kind):
    """f = f_eq(rho_t, u(x)) + (f_eq(x + c) - f(x + c)) leaves a uniform
    equilibrium flow at rho_t = 1 unchanged (the RF correction needs u_x < 0)."""
    u0 = 0.04
    flags, _, bc = _channel(12, 6, 6, u_in=u0, obstacle=False, walls=False, outlet_kind=kind)
    o = OracleLattice(make_desc(flags, bc, periodic=(0, 1, 0), collision=abi.COLLISION_CUM2015,
                                storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=0, tau=0.6, nu=0.1 / 3))
    u = np.float64(np.float32(u0))
    o.init_equilibrium(None, np.full(flags.shape, u), None, None)
    o.step(20)
    expect = feq(1.0, (u, 0.0, 0.0))
    assert np.abs(o.populations() - expect[:, None, None, None]).max() < 1e-15


def test_density_equilibrium_outlet_writes_the_outgoing_equilibrium_into_the_incoming_slot():
    """OutletDensityEquilibrium.h:64-71: slot c_rev[v] of every outgoing v
    (n.c > 0) holds f_eq(v; rho_t, u(x)) after the step.  At rest the rule is
    a fixed point; in a flow it is not (f_eq(v) != f_eq(c_rev v))."""
    c, crev, _ = tables(0)
    u0 = 0.04
    flags, _, bc = _channel(12, 6, 6, u_in=u0, obstacle=False, walls=False, outlet_kind=abi.BC_OUTLET_EQUILIBRIUM)
    o = OracleLattice(make_desc(flags, bc, periodic=(0, 1, 0), collision=abi.COLLISION_CUM2015,
                                storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=0, tau=0.6, nu=0.1 / 3))
    u = np.float64(np.float32(u0))
    o.init_equilibrium(None, np.full(flags.shape, u), None, None)
    o.step(1)
    f = o.populations()
    eq = feq(1.0, (u, 0.0, 0.0))
    out_x = flags.shape[2] - 1
    outgoing = [v for v in range(27) if c[v][0] > 0]
    for v in outgoing:
        assert np.allclose(f[crev[v], :, :, out_x], eq[v], atol=1e-15, rtol=0)
        assert abs(eq[v] - eq[crev[v]]) > 1e-4
    flags2, _, bc2 = _channel(12, 6, 6, u_in=0.0, obstacle=False, walls=False,
                              outlet_kind=abi.BC_OUTLET_EQUILIBRIUM)
    o2 = OracleLattice(make_desc(flags2, bc2, periodic=(0, 1, 0), collision=abi.COLLISION_CUM2015,
                                 storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=0, tau=0.6, nu=0.1 / 3))
    o2.init_equilibrium(None, 0.0, None, None)
    o2.step(5)
    assert np.abs(o2.populations() - feq(1.0, (0, 0, 0))[:, None, None, None]).max() < 1e-15


def test_outlet_omega_damping_is_a_ramp_over_the_last_tenth():
    """outletOmega (OutletNeighbourEquilibriumOmega.h:18-79): omega = 1/tau
    outside the last ceil(0.1 n) cells of the outlet normal's axis and decreases
    linearly inside them, by (high - low) / ceil(0.1 n) per cell, from 0 at the
    first damped cell."""
    nx, tau, lo, hi = 25, 0.6, 0.0, float(np.float32(0.05))   # float constants, as the reference's
    flags, _, bc = _channel(nx, 6, 6, obstacle=False, walls=False, outlet_kind=abi.BC_OUTLET_NEIGHBOUR_EQ)
    o = OracleLattice(make_desc(flags, bc, periodic=(0, 1, 0), collision=abi.COLLISION_CUM2015,
                                storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=0, tau=tau, nu=0.1 / 3,
                                sponge=(lo, hi)))
    o.init_equilibrium(None, 0.02, None, None)
    o.step(1)
    w = o.omega()
    m = int(np.ceil(nx * np.float32(0.1)))   # 3 cells
    x = np.arange(nx)
    ramp = np.where(x >= nx - m, (hi - lo) / m * (x - (nx - m)), 0.0)
    fluid = flags != 0
    expect = np.broadcast_to(1 / tau - ramp, flags.shape)
    assert m == 3
    assert np.allclose(w[fluid], expect[fluid], rtol=0, atol=1e-15)


def test_zou_he_inlet_keeps_a_uniform_inflow():
    """InletVelocityZouHe: incoming = (f* - f_eq)(outgoing) + f_eq(incoming; rho, u_in).
    A uniform equilibrium flow at the inlet velocity is a fixed point up to
    the float evaluation of the inlet equilibrium's polynomial (~1e-8)."""
    u0 = 0.04
    flags, _, bc = _channel(12, 6, 6, u_in=u0, obstacle=False, walls=False, inlet_kind=abi.BC_INLET_ZOU_HE)
    o = OracleLattice(make_desc(flags, bc, periodic=(0, 1, 0), collision=abi.COLLISION_CUM2015,
                                storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=0, tau=0.6, nu=0.1 / 3))
    u = np.float64(np.float32(u0))
    o.init_equilibrium(None, np.full(flags.shape, u), None, None)
    o.step(20)
    expect = feq(1.0, (u, 0.0, 0.0))
    err = np.abs(o.populations() - expect[:, None, None, None]).max()
    assert 0 < err < 1e-8
    # a mismatched inflow is pulled towards the inlet velocity at the inlet cells
    o2 = OracleLattice(make_desc(flags, bc, periodic=(0, 1, 0), collision=abi.COLLISION_CUM2015,
                                 storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=0, tau=0.6, nu=0.1 / 3))
    o2.init_equilibrium(None, 0.0, None, None)
    o2.step(30)
    inlet = flags == -1
    assert np.all(o2.macroscopic()[1][inlet] > 0.25 * u0)


# --- BounceBackWallHalfMesh -----------------------------------------------

def _sparse_obstacles(n=12):
    """A periodic box with a few solid cells and a small solid block: their
    26-neighbours carry flags 2 (one solid neighbour) and higher."""
    from oracle import mesher_oracle as mo
    g = np.ones((n, n, n), np.int32)
    g[3, 4, 5] = 0
    g[8, 2, 9] = 0
    g[6:8, 7:9, 2:4] = 0
    flags, walls = mo.meshing_boundary_wall(g)
    walls = mo.compile_wall(flags, walls)
    return flags, walls


def test_wall_mesh_policy_is_the_list_rule_on_flag_two_cells():
    """BounceBackWallHalfMesh (BounceBackWallHalfMesh.h:38-50) applies the
    half-way rule of BounceBackWallHalf to every flag-2 cell: it equals the
    list policy with the list replaced by exactly those cells."""
    from helpers import bc_nodes
    flags, walls = _sparse_obstacles()
    assert (flags == 2).sum() > 20 and len(walls) > 0
    listed = bc_nodes([dict(kind=abi.BC_WALL, x=w[0], y=w[1], z=w[2], regular=w[3]) for w in walls])
    flag2 = [dict(kind=abi.BC_WALL, x=int(x), y=int(y), z=int(z), regular=1)
             for z, y, x in zip(*np.nonzero(flags == 2))]
    u = random_velocity_field(flags.shape, 0.03, seed=13)
    u = [np.where(flags == 0, 0.0, a) for a in u]
    kw = dict(collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL,
              compat=abi.COMPAT_REFERENCE, periodic=(1, 1, 1), tau=0.56)
    out = []
    for bc, mesh in ((listed, True), (bc_nodes(flag2), False), (listed, False)):
        o = OracleLattice(make_desc(flags, bc, wall_mesh=mesh, **kw))
        o.init_equilibrium(None, *u)
        o.step(12)
        out.append(o.populations())
    assert np.array_equal(out[0], out[1])
    assert not np.array_equal(out[0], out[2])   # the two policies act on different cells
