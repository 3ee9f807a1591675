"""Stochastic-rounding deviation storage (F16DEV_SR, BF16DEV_SR) on the GPU.

The collide-stream step stores g = f - w*rho0 rounded stochastically: each
stored value is one of the two neighbours of g in the 16-bit format, the
upper one with probability (g - lo) / (hi - lo), so E[stored] = g.  The
format has no counterpart in the reference (its populations are FP32/FP64),
so it is pinned by these properties rather than by the oracle:

- on a uniform lattice at rest every cell computes the same FP32 g, so across
  cells the stored values of one population take at most two adjacent 16-bit
  values, whose frequencies match g's position between them (binomial bound);
  round-to-nearest stores a single value everywhere;
- runs are reproducible (the random bits hash the cell and the step counter);
- on a Taylor-Green flow the velocity error against FP64 storage stays well
  below round-to-nearest's, whose rounding error repeats from step to step on
  the slowly changing field and adds up.
"""
import numpy as np
import pytest

from helpers import random_velocity_field, rel_l2
from lb_amd import Lattice, abi, make_desc
from lb_amd.fields import taylor_green

pytestmark = pytest.mark.gpu

N_REST = 32
DELTA = 0.0137


def _lattice(storage, n, tau=0.6, **kw):
    flags = np.ones((n, n, n), np.int32)
    return Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2015, storage=storage, arith=abi.ARITH_FAST,
                             compat=0, periodic=(1, 1, 1), tau=tau, **kw))


def _bf16_rn(x32):
    b = x32.astype(np.float32).view(np.uint32).astype(np.uint64)
    return ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)


def _bf16_value(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _stored(f, w32, fmt):
    """The 16-bit values behind populations f (FP64 of (float)h + w): the
    format's spacing is far above FP32's at these magnitudes, so rounding
    f - w to the format recovers h exactly."""
    d = (f - w32[:, None]).astype(np.float32)
    if fmt == "f16":
        return d.astype(np.float16).astype(np.float64)
    return _bf16_value(_bf16_rn(d))


def _next_up(v, fmt):
    if fmt == "f16":
        return float(np.nextafter(np.float16(v), np.float16(np.inf)))
    b = int(_bf16_rn(np.array([v], np.float32))[0])
    return float(_bf16_value(np.array([b + 1 if v >= 0 else b - 1], np.uint16))[0])


@pytest.mark.parametrize("sr,rn,fmt", [(abi.STORAGE_F16DEV_SR, abi.STORAGE_F16DEV, "f16"),
                                       (abi.STORAGE_BF16DEV_SR, abi.STORAGE_BF16DEV, "bf16")],
                         ids=["f16", "bf16"])
def test_rest_state_rounds_to_neighbours_without_bias(gpu, sr, rn, fmt):
    n = N_REST
    shape = (n, n, n)
    ref = _lattice(abi.STORAGE_F64, 2)
    ref.init_equilibrium(None, None, None, None)
    w = ref.populations()[:, 0, 0, 0]               # rho = 1, u = 0: f = w
    w32 = w.astype(np.float32).astype(np.float64)   # the kernels' w * rho0 in FP32
    # tau = 1 relaxes every moment to equilibrium: from a uniform rest state
    # (set_populations, so the step takes its moments from the populations)
    # the post-collision state is w * rho_c with rho_c the stored density
    a, b = _lattice(sr, n, tau=1.0), _lattice(rn, n, tau=1.0)
    f_set = np.broadcast_to((w * (1.0 + DELTA))[:, None, None, None], (abi.Q, *shape))
    for lat in (a, b):
        lat.set_populations(f_set)
    f0 = a.populations().reshape(abi.Q, -1)
    rho_c = f0[:, 0].sum()
    a.step(1)
    b.step(1)
    ga = _stored(a.populations().reshape(abi.Q, -1), w32, fmt)
    gb = _stored(b.populations().reshape(abi.Q, -1), w32, fmt)
    ncell = ga.shape[1]
    worst = 0.0
    for q in range(abi.Q):
        g = w[q] * (rho_c - 1.0)                    # post-collision deviation at rest
        assert np.unique(gb[q]).size == 1           # nearest rounding: one value everywhere
        vals = np.unique(ga[q])
        assert vals.size <= 2, (q, vals)
        lo = vals[0] if g >= 0 else vals[-1]        # the neighbour nearer zero
        hi = _next_up(lo, fmt) if g >= 0 else -_next_up(-lo, fmt)
        if vals.size == 2:
            assert {vals[0], vals[1]} == {lo, hi}, (q, vals, lo, hi)
        p = (g - lo) / (hi - lo)                    # expected share of the outer neighbour
        p_hat = np.count_nonzero(ga[q] == hi) / ncell
        sigma = np.sqrt(max(p * (1 - p), 1e-4) / ncell)
        tol = 5 * sigma + 0.08                      # 0.08: FP32 rounding of the kernel's g against g
        worst = max(worst, abs(p_hat - p) / tol)
        assert -0.08 <= p <= 1.08 and abs(p_hat - p) <= tol, (q, p, p_hat, tol)
    print(f"[sr rest {fmt}] worst |p_hat - p| / tol = {worst:.3f}", flush=True)


@pytest.mark.parametrize("storage", [abi.STORAGE_F16DEV_SR, abi.STORAGE_BF16DEV_SR], ids=["f16", "bf16"])
def test_stochastic_rounding_is_reproducible(gpu, storage):
    shape = (24, 24, 24)
    u = random_velocity_field(shape, 0.04, seed=31)
    runs = []
    for st in (storage, storage, storage - 2):      # twice with SR, once with nearest
        lat = _lattice(st, 24, tau=0.52, closure=abi.CLOSURE_SMAGORINSKY, c_les=0.17)
        lat.init_equilibrium(None, *u)
        lat.step(10)
        runs.append(lat.populations())
    assert np.array_equal(runs[0], runs[1])
    assert not np.array_equal(runs[0], runs[2])
    assert rel_l2(runs[0], runs[2]) <= 1e-3


def test_taylor_green_error_below_nearest_rounding(gpu):
    """64^3 Taylor-Green, u0 = 0.01, Re 100, 1000 steps (about one eddy
    time, still laminar) against FP64 storage: a slowly changing field, on
    which round-to-nearest repeats its error every step."""
    n, u0 = 64, 0.01
    nu = u0 * (n / (2 * np.pi)) / 100.0
    rho, ux, uy, uz = taylor_green(n, u0)
    err = {}
    lats = {}
    for st in (abi.STORAGE_F64, abi.STORAGE_F16DEV, abi.STORAGE_F16DEV_SR, abi.STORAGE_BF16DEV,
               abi.STORAGE_BF16DEV_SR):
        lat = _lattice(st, n, tau=3 * nu + 0.5)
        lat.init_equilibrium(rho, ux, uy, uz)
        lat.step(1000)
        lats[st] = np.stack(lat.macroscopic()[1:])
    for st in lats:
        if st != abi.STORAGE_F64:
            err[abi.STORAGE_NAMES[st]] = rel_l2(lats[st], lats[abi.STORAGE_F64])
    print(f"[tgv 64^3 Re 100, 1000 steps] relL2(u) vs f64: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    assert err["f16dev_sr"] <= 1e-3
    assert err["f16dev_sr"] <= err["f16dev"] / 3
    assert err["bf16dev_sr"] <= err["bf16dev"] / 3
