"""Learned closure, CPU side: the FP64 network oracle against finite
differences, the feature definitions against linear algebra and against the
product header compiled for the host, and the SGS target against a field
whose sub-grid stress is an exact eddy viscosity."""
import ctypes

import numpy as np

from helpers import host_collide_lib, oracle_lib
from lb_amd import abi, closure
from oracle import mlp_oracle as mo


def _random_problem(n, seed):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed)
    p[mo.OFF_W3:mo.OFF_B3] = 0.3 * rng.standard_normal(mo.NH)
    p[mo.OFF_B1:mo.OFF_W2] = 0.1 * rng.standard_normal(mo.NH)
    p[mo.OFF_B3] = 0.2
    X = rng.standard_normal((n, mo.NF))
    y = 0.17 * (1 + 0.3 * rng.standard_normal(n))
    w = rng.uniform(0.5, 2.0, n)
    return p, X, y, w


def test_oracle_backward_matches_central_differences():
    p, X, y, w = _random_problem(40, 1)
    c = 0.17
    _, g = mo.loss_grad(p, X, y, w, c)
    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.choice(np.arange(mo.OFF_W1, mo.OFF_B1), 8), rng.choice(np.arange(mo.OFF_B1, mo.OFF_W2), 4),
                          rng.choice(np.arange(mo.OFF_W2, mo.OFF_B2), 8), rng.choice(np.arange(mo.OFF_B2, mo.OFF_W3), 4),
                          rng.choice(np.arange(mo.OFF_W3, mo.OFF_B3), 4), [mo.OFF_B3]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_grad(pp, X, y, w, c)[0] - mo.loss_grad(pm, X, y, w, c)[0]) / (2 * h) / len(y)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])), (i, fd, g[i])


def test_head_is_smagorinsky_at_zero():
    p = mo.init_params(3)
    C, _ = mo.forward(p, np.zeros((4, mo.NF)), 0.17)
    assert np.allclose(C, 0.17, rtol=1e-14)


def _strain(rng, n):
    return rng.standard_normal((6, n)) * 10 ** rng.uniform(-4, -2, n)


def test_strain_features_against_linear_algebra():
    rng = np.random.default_rng(4)
    S = _strain(rng, 200)
    x = closure.strain_features(S, 1.0 + 0.01 * rng.standard_normal(200), 0.51)
    M = np.empty((200, 3, 3))
    M[:, 0, 0], M[:, 1, 1], M[:, 2, 2] = S[0], S[1], S[2]
    M[:, 0, 1] = M[:, 1, 0] = S[3]
    M[:, 0, 2] = M[:, 2, 0] = S[4]
    M[:, 1, 2] = M[:, 2, 1] = S[5]
    nrm = np.sqrt((M * M).sum((1, 2)))
    assert np.allclose(10 ** x[0], nrm, rtol=1e-12)
    D = M - np.trace(M, axis1=1, axis2=2)[:, None, None] / 3 * np.eye(3)
    lam = np.linalg.eigvalsh(D)
    dn = np.sqrt((D * D).sum((1, 2)))
    assert np.allclose(x[8], np.sqrt(6) * (lam ** 3).sum(1) / dn ** 3, atol=1e-10)
    assert np.all(np.abs(x[8]) <= 1 + 1e-12)
    assert np.allclose(x[11], dn / nrm, rtol=1e-12)
    assert np.allclose(x[1] ** 2 + x[2] ** 2 + x[3] ** 2 + 2 * (x[4] ** 2 + x[5] ** 2 + x[6] ** 2), 1.0, rtol=1e-12)
    assert np.allclose(x[10], np.log10(nrm / ((0.51 - 0.5) / 3)), rtol=1e-12)


def test_strain_features_match_product_header():
    hc = host_collide_lib()
    rng = np.random.default_rng(5)
    S = _strain(rng, 50)
    rho = 1.0 + 0.01 * rng.standard_normal(50)
    x = closure.strain_features(S, rho, 0.505)
    out = np.empty(12)
    for i in range(50):
        s6 = np.ascontiguousarray(S[:, i])
        hc.host_strain_features(1, abi.ptr(s6), rho[i], 0.505, abi.ptr(out))
        assert np.allclose(out, x[:, i], rtol=1e-13, atol=1e-13)
        hc.host_strain_features(0, abi.ptr(s6), rho[i], 0.505, abi.ptr(out))
        assert np.allclose(out, x[:, i], rtol=2e-4, atol=2e-5)


def test_cell_features_recover_the_strain_behind_the_stress():
    """A cell f = f_eq(rho, u) + delta whose non-equilibrium stress is
    Pi = -(2/3) rho tau_s S for a chosen S (tau_s the Smagorinsky rate of
    that Pi) gives back S through its features."""
    hc = host_collide_lib()
    lib = oracle_lib()
    rng = np.random.default_rng(6)
    c = np.zeros((27, 3), np.int32)
    cr = np.zeros(27, np.int32)
    wt = np.zeros(27)
    lib.orc_tables(0, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), cr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                   abi.ptr(wt))
    tau0, cles, rho = 0.51, 0.17, 1.02
    u = np.array([0.03, -0.01, 0.02])
    Pi = rng.standard_normal(6) * 1e-4
    Pi[:3] -= Pi[:3].mean()   # traceless stress
    fe = np.empty(27)
    lib.orc_cell_feq(0, 1, rho, *u, abi.ptr(fe))
    # delta_q = w_q * 4.5 * (c c - I/3) : Pi  carries Pi and no mass/momentum
    Pm = np.array([[Pi[0], Pi[3], Pi[4]], [Pi[3], Pi[1], Pi[5]], [Pi[4], Pi[5], Pi[2]]])
    d = np.array([wt[q] * 4.5 * np.sum((np.outer(c[q], c[q]) - np.eye(3) / 3) * Pm) for q in range(27)])
    f = np.ascontiguousarray(fe + d)
    x = np.empty(12)
    pin = ctypes.c_double()
    om = ctypes.c_double()
    hc.host_cell_features(abi.ptr(f), tau0, cles, abi.ptr(x), ctypes.byref(pin), ctypes.byref(om))
    npi = np.sqrt((Pm * Pm).sum())
    assert abs(pin.value - npi / rho) < 1e-12 * npi
    tau_s = 0.5 * tau0 + 0.5 * np.sqrt(tau0 ** 2 + 18 * cles * npi / rho)
    assert abs(om.value - 1 / tau_s) < 1e-14
    S = -Pm / ((2 / 3) * rho * tau_s)
    nS = np.sqrt((S * S).sum())
    assert abs(10 ** x[0] - nS) < 1e-10 * nS
    assert np.allclose(x[1:7] * nS, [S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]], atol=1e-10 * nS)


def test_sgs_target_recovers_an_exact_eddy_viscosity():
    """tau_dev = -2 nu S_dev exactly gives nu_t* = nu and C* = nu / |S|."""
    rng = np.random.default_rng(7)
    n = 12
    k = 2 * np.pi * np.arange(n) / n
    z, y, x = np.meshgrid(k, k, k, indexing="ij")
    u = 0.02 * np.sin(x + 2 * y) + 0.01 * np.cos(z)
    v = 0.015 * np.cos(x - z) + 0.005 * rng.standard_normal((n, n, n))
    w = 0.01 * np.sin(y + z)
    nu = 3e-3
    S = closure.strain_rate(u, v, w)
    tr = (S[0] + S[1] + S[2]) / 3
    t = [-2 * nu * (S[0] - tr), -2 * nu * (S[1] - tr), -2 * nu * (S[2] - tr), -2 * nu * S[3], -2 * nu * S[4],
         -2 * nu * S[5]]
    filt = np.stack([np.ones_like(u), u, v, w, t[0] + u * u, t[1] + v * v, t[2] + w * w, t[3] + u * v, t[4] + u * w,
                     t[5] + v * w])
    _, C, nS = closure.sgs_targets(filt)
    assert np.allclose(C * nS, nu, rtol=1e-9)


def test_normalisation_standardises():
    rng = np.random.default_rng(8)
    X = rng.standard_normal((5000, 12)) * np.arange(1, 13) + np.arange(12)
    norm = closure.fit_normalisation(X)
    Z = closure.normalise(X, norm)
    assert np.allclose(Z.mean(0), 0, atol=1e-5) and np.allclose(Z.std(0), 1, atol=1e-4)
    assert np.all(norm[2] < -2.5) and np.all(norm[3] > 2.5)
