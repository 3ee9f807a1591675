"""Learned closure on the GPU: training gradients and inference against the
FP64 network oracle, the fused collide-stream kernel against the stand-alone
inference kernel, the Smagorinsky fallback, and the box filter."""
import numpy as np
import pytest

from helpers import random_velocity_field, rel_l2
from lb_amd import Lattice, abi, closure, make_desc
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu

C_LES = 0.17


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed)
    p[mo.OFF_W3:mo.OFF_B3] = 0.3 * rng.standard_normal(mo.NH)
    p[mo.OFF_B1:mo.OFF_W2] = 0.1 * rng.standard_normal(mo.NH)
    p[mo.OFF_B2:mo.OFF_W3] = 0.1 * rng.standard_normal(mo.NH)
    p[mo.OFF_B3] = 0.2
    X = rng.standard_normal((n, mo.NF)).astype(np.float32)
    y = (C_LES * (1 + 0.3 * rng.standard_normal(n))).astype(np.float32)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    return p.astype(np.float32), X, y, w


@pytest.mark.parametrize("bf16,tol", [(False, 1e-5), (True, 2e-2)], ids=["fp32", "bf16"])
def test_gradient_matches_oracle(gpu, bf16, tol):
    """FP32 products: the data flow matches the analytic backward pass to
    FP32 rounding.  BF16 matrix-core operands: within 2e-2 relative L2."""
    p, X, y, w = _problem(2048, 11)
    m = closure.Mlp(params=p)
    m.set_data(X, y, w)
    loss, g = m.gradient(0, 2048, C_LES, bf16=bf16)
    lo, go = mo.loss_grad(p.astype(np.float64), X.astype(np.float64), y, w, C_LES)
    eg = rel_l2(g, go)
    el = abs(loss - lo) / lo
    blocks = {"W1": (mo.OFF_W1, mo.OFF_B1), "W2": (mo.OFF_W2, mo.OFF_B2), "w3": (mo.OFF_W3, mo.OFF_B3)}
    per = {k: rel_l2(g[a:b], go[a:b]) for k, (a, b) in blocks.items()}
    print(f"[mlp gradient bf16={bf16}] relL2 grad={eg:.3e} loss={el:.3e} per-block={per}", flush=True)
    assert eg <= tol and el <= (1e-6 if not bf16 else 1e-2)
    assert all(v <= tol for v in per.values())


def test_gradient_sub_batches_add_up(gpu):
    p, X, y, w = _problem(1024, 12)
    m = closure.Mlp(params=p)
    m.set_data(X, y, w)
    l_all, g_all = m.gradient(0, 1024, C_LES, bf16=False)
    l_a, g_a = m.gradient(0, 512, C_LES, bf16=False)
    l_b, g_b = m.gradient(512, 512, C_LES, bf16=False)
    assert abs(l_a + l_b - l_all) <= 1e-9 * l_all
    assert rel_l2(0.5 * (g_a + g_b), g_all) <= 1e-6


def test_inference_matches_oracle(gpu):
    p, X, _, _ = _problem(4096, 13)
    m = closure.Mlp(params=p)
    C = m.forward(X, C_LES)
    Co, _ = mo.forward(p.astype(np.float64), X.astype(np.float64), C_LES, round_bf16=True)
    err = np.abs(C - Co) / np.abs(Co)
    print(f"[mlp inference] max rel err {err.max():.3e} mean {err.mean():.3e}", flush=True)
    assert err.max() <= 1e-2 and err.mean() <= 1e-3


def test_training_fits_a_smooth_target(gpu):
    rng = np.random.default_rng(14)
    n = 65536
    X = rng.standard_normal((n, 12)).astype(np.float32)
    y = (C_LES * (1 + 0.5 * np.tanh(X[:, 0]) + 0.2 * X[:, 1] * X[:, 2])).clip(0.01).astype(np.float32)
    m = closure.Mlp(seed=3)
    m.set_data(X, y, None)
    loss = m.train(600, 4096, 3e-3, C_LES, seed=5)
    print(f"[mlp training] loss {loss[:20].mean():.3e} -> {loss[-20:].mean():.3e}", flush=True)
    assert np.all(np.isfinite(loss))
    assert loss[-20:].mean() < 0.1 * loss[:20].mean()
    # the same steps from the same state give the same parameters
    m2 = closure.Mlp(seed=3)
    m2.set_data(X, y, None)
    m2.train(600, 4096, 3e-3, C_LES, seed=5)
    assert np.array_equal(m.params(), m2.params())


def _turbulent_lattice(n, closure_kind, c_les=C_LES, storage=abi.STORAGE_F32, seed=21):
    flags = np.ones((n, n, n), np.int32)
    lat = Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2015, closure=closure_kind, storage=storage,
                            arith=abi.ARITH_FAST, compat=0, periodic=(1, 1, 1), tau=0.503, c_les=c_les))
    u = random_velocity_field(flags.shape, 0.05, seed=seed, kmax=4)
    return lat, (None, u[0], u[1], u[2])


def _open_norm():
    return np.stack([np.zeros(12), np.ones(12), np.full(12, -1e30), np.full(12, 1e30)]).astype(np.float32)


def test_fused_inference_equals_standalone_kernel(gpu):
    p, _, _, _ = _problem(16, 15)
    lat, init = _turbulent_lattice(24, abi.CLOSURE_MLP)
    lat.set_closure(p, _open_norm())
    lat.init_equilibrium(*init)
    lat.step(5)
    raw, _, _ = lat.closure_features()
    X = raw.reshape(12, -1).T
    norm = closure.fit_normalisation(X, q=(0.05, 0.95))
    lat.set_closure(p, norm)
    feats, env, coef = lat.closure_features()
    Xn = feats.reshape(12, -1).T
    assert np.allclose(Xn, closure.normalise(X, norm), rtol=1e-5, atol=1e-5)
    m = closure.Mlp(params=p)
    C = m.forward(Xn, C_LES)
    inside = env.ravel()
    print(f"[fused vs standalone] in envelope {inside.mean():.3f}", flush=True)
    assert 0.2 < inside.mean() < 1.0
    assert np.array_equal(coef.ravel()[inside].astype(np.float32), C[inside])
    assert np.all(coef.ravel()[~inside] == np.float32(C_LES))
    # the step runs with these coefficients: omega matches the rate formula
    om = lat.omega().ravel()
    _, _, _, _ = lat.macroscopic()
    lat.step(3)
    assert np.all(np.isfinite(lat.macroscopic()[1]))
    assert np.all((om > 0) & (om < 2))


def test_out_of_envelope_equals_smagorinsky(gpu):
    """Outside the envelope the rate is the Smagorinsky formula bit for bit
    (same state -> same omega).  Over steps the two kernels stay within
    FP32 rounding: the shared collision source is compiled into two kernels
    whose multiply-add fusion may differ."""
    p, _, _, _ = _problem(16, 16)
    norm = _open_norm()
    norm[2], norm[3] = 1.0, -1.0   # empty envelope: every cell falls back
    a, init = _turbulent_lattice(24, abi.CLOSURE_MLP)
    a.set_closure(p, norm)
    b, _ = _turbulent_lattice(24, abi.CLOSURE_SMAGORINSKY)
    a.init_equilibrium(*init)
    b.init_equilibrium(*init)
    assert np.array_equal(a.omega(), b.omega())
    a.step(10)
    b.step(10)
    fa, fb = a.populations(), b.populations()
    print(f"[fallback] max|df| {np.abs(fa - fb).max():.3e}", flush=True)
    assert rel_l2(fa, fb) <= 1e-6
    _, env, coef = a.closure_features()
    assert not env.any() and np.all(coef == np.float32(C_LES))
    b.set_populations(fa)
    assert np.array_equal(a.omega(), b.omega())


def test_constant_network_equals_smagorinsky_with_its_coefficient(gpu):
    p = np.zeros(mo.NPARAM, np.float32)
    p[mo.OFF_B3] = 0.7
    C0 = float(np.float32(C_LES * np.log1p(np.exp(0.7)) / np.log(2.0)))
    a, init = _turbulent_lattice(24, abi.CLOSURE_MLP)
    a.set_closure(p, _open_norm())
    b, _ = _turbulent_lattice(24, abi.CLOSURE_SMAGORINSKY, c_les=C0)
    a.init_equilibrium(*init)
    b.init_equilibrium(*init)
    a.step(10)
    b.step(10)
    ua, ub = a.macroscopic(), b.macroscopic()
    e = rel_l2(np.stack(ua[1:]), np.stack(ub[1:]))
    print(f"[constant network] relL2(u) {e:.3e}", flush=True)
    assert e <= 1e-5


def test_box_filter_matches_host_average(gpu):
    n = 16
    flags = np.ones((n, n, n), np.int32)
    lat = Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F64,
                            arith=abi.ARITH_LITERAL, compat=0, periodic=(1, 1, 1), tau=0.52))
    u = random_velocity_field(flags.shape, 0.05, seed=22, kmax=3)
    lat.init_equilibrium(1.0 + 0.01 * u[0], u[0], u[1], u[2])
    out0 = lat.filter(4)
    lat.step(3)
    rho, ux, uy, uz = lat.macroscopic()
    out = lat.filter(2)

    def box(a, f):
        m = n // f
        return a.reshape(m, f, m, f, m, f).mean((1, 3, 5))

    ref = [box(a, 2) for a in (rho, ux, uy, uz, ux * ux, uy * uy, uz * uz, ux * uy, ux * uz, uy * uz)]
    for k in range(10):
        assert np.allclose(out[k], ref[k], rtol=1e-12, atol=1e-15), k
    # before the first step the filter sees the initial fields
    assert np.allclose(out0[1], box(u[0], 4), rtol=1e-12, atol=1e-15)
    # a constant field filters to itself
    c = Lattice(make_desc(flags, [], storage=abi.STORAGE_F64, periodic=(1, 1, 1), tau=0.52))
    c.init_equilibrium(None, np.full(flags.shape, 0.03), None, None)
    c.step(1)
    assert np.allclose(c.filter(4)[1], 0.03, rtol=1e-12)


def test_refresh_holds_the_network_coefficient(gpu):
    """With a refresh interval of 3 the network's coefficients from the
    evaluating step are held for the next two steps, then re-evaluated."""
    p, _, _, _ = _problem(16, 17)
    lat, init = _turbulent_lattice(24, abi.CLOSURE_MLP)
    lat.set_closure(p, _open_norm())
    lat.set_closure_refresh(3)
    lat.init_equilibrium(*init)
    lat.step(3)                         # one full cycle; the next step evaluates
    _, _, c_fresh = lat.closure_features()
    lat.step(1)                         # evaluates the network on this state and holds it
    _, _, c_held1 = lat.closure_features()
    lat.step(1)
    _, _, c_held2 = lat.closure_features()
    assert np.array_equal(c_held1, c_fresh) and np.array_equal(c_held2, c_fresh)
    lat.step(1)                         # cycle complete: the next step re-evaluates
    _, _, c_new = lat.closure_features()
    assert not np.array_equal(c_new, c_fresh)
    assert np.all(np.isfinite(lat.macroscopic()[1]))


@pytest.mark.parametrize("inlet_kind", ["moving_wall", "zou_he"])
@pytest.mark.parametrize("refresh", [1, 10])
def test_fallback_on_a_lattice_with_boundaries_equals_smagorinsky(gpu, inlet_kind, refresh):
    """The flagged instantiations of the network step and of the held-
    coefficient step (walls, a moving-wall or Zou-He inlet, an interpolated
    outlet, ghost layers): with an empty envelope every cell falls back, and
    the run equals the Smagorinsky run on the same lattice to FP32 rounding."""
    from helpers import bc_nodes
    from oracle import mesher_oracle as mesh

    kind = mesh.BC_INLET_MOVING_WALL if inlet_kind == "moving_wall" else mesh.BC_INLET_ZOU_HE
    flags, bcd = mesh.channel(32, 16, 12, inlet_kind=kind)
    bc = bc_nodes(bcd)
    assert (flags == 0).any() and (flags == -1).any() and (flags == -2).any() and (flags == -3).any()
    p, _, _, _ = _problem(16, 18)
    norm = _open_norm()
    norm[2], norm[3] = 1.0, -1.0
    kw = dict(collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F32, arith=abi.ARITH_FAST, compat=0,
              tau=0.53, nu=0.01, c_les=C_LES)
    a = Lattice(make_desc(flags, bc, closure=abi.CLOSURE_MLP, **kw))
    a.set_closure(p, norm)
    a.set_closure_refresh(refresh)
    b = Lattice(make_desc(flags, bc, closure=abi.CLOSURE_SMAGORINSKY, **kw))
    rng = np.random.default_rng(19)
    solid = flags == 0
    ux = np.where(solid, 0.0, 0.04 + 0.004 * rng.standard_normal(flags.shape))
    uy = np.where(solid, 0.0, 0.004 * rng.standard_normal(flags.shape))
    for L in (a, b):
        L.init_equilibrium(None, ux, uy, None)
        L.step(30)
    fa, fb = a.populations(), b.populations()
    e = rel_l2(fa, fb)
    print(f"[fallback with boundaries, {inlet_kind}, refresh {refresh}] relL2 {e:.3e}", flush=True)
    assert np.all(np.isfinite(fa)) and e <= 1e-6
    _, env, coef = a.closure_features()
    assert not env.any()
