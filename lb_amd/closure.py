"""Learned sub-grid closure: data, training and evaluation (host side).

The network and its kernels live in the HIP core (include/lbm_closure.h);
this module prepares what they consume and runs the experiment around them:

* training data from the solver's own fine-grid runs: the fine lattice is
  box-filtered on the GPU (lbm_filter) and, on the coarse grid, the resolved
  strain S = sym grad u (central differences, periodic), the sub-grid stress
  tau = <u u> - <u><u> and the least-squares eddy viscosity
  nu_t* = -tau_dev : S / (2 S : S) give the target coefficient C* = nu_t*/|S|_F
  (the coefficient of the reference's Smagorinsky rate formula,
  OmegaLES.h:103-131, is exactly nu_t / |S|_F);
* features (the numpy mirror of lb_amd/csrc/features.hpp), normalisation and
  the training envelope;
* training on the GPU through lbm_mlp_train;
* evaluation: coarse LES runs with the Smagorinsky rate and with the network,
  started from the filtered fine field, against the filtered fine run's
  kinetic-energy decay (dissipation) and spectrum.
"""
from __future__ import annotations

import ctypes
import json
import math
import time

import numpy as np

from . import abi
from .core import Lattice, _check, make_desc
from .fields import decaying_turbulence, energy_spectrum

NF = abi.MLP_FEATURES
NPARAM = abi.MLP_PARAMS


class Mlp:
    """One closure network on the GPU (lbm_mlp_*)."""

    def __init__(self, params=None, seed=1):
        self._lib = abi.load_core()
        h = ctypes.c_void_p()
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        _check(self._lib.lbm_mlp_create(abi.ptr(p, ctypes.c_float), int(seed), ctypes.byref(h)), "lbm_mlp_create")
        self._h = h
        self._data = None

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lbm_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_data(self, X, y, w=None):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        w = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
        _check(self._lib.lbm_mlp_set_data(self._h, abi.ptr(X, ctypes.c_float), abi.ptr(y, ctypes.c_float),
                                          abi.ptr(w, ctypes.c_float), X.shape[0]), "lbm_mlp_set_data")
        self._data = X.shape[0]

    def gradient(self, first, count, c_les, bf16=True):
        g = np.empty(NPARAM, np.float32)
        loss = ctypes.c_double()
        _check(self._lib.lbm_mlp_gradient(self._h, int(first), int(count), float(c_les), int(bool(bf16)),
                                          ctypes.byref(loss), abi.ptr(g, ctypes.c_float)), "lbm_mlp_gradient")
        return loss.value, g

    def train(self, steps, batch, lr, c_les, seed=1):
        loss = np.empty(steps)
        _check(self._lib.lbm_mlp_train(self._h, int(steps), int(batch), float(lr), float(c_les), int(seed),
                                       abi.ptr(loss)), "lbm_mlp_train")
        return loss

    def forward(self, X, c_les):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.empty(X.shape[0], np.float32)
        _check(self._lib.lbm_mlp_forward(self._h, abi.ptr(X, ctypes.c_float), X.shape[0], float(c_les),
                                         abi.ptr(out, ctypes.c_float)), "lbm_mlp_forward")
        return out

    def params(self):
        p = np.empty(NPARAM, np.float32)
        _check(self._lib.lbm_mlp_get_params(self._h, abi.ptr(p, ctypes.c_float)), "lbm_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(self._lib.lbm_mlp_set_params(self._h, abi.ptr(p, ctypes.c_float)), "lbm_mlp_set_params")


# --- features (mirror of features.hpp) -------------------------------------

def strain_features(S, rho, tau0):
    """Features [12, ...] from strain components S = (xx, yy, zz, xy, xz, yz)."""
    sxx, syy, szz, sxy, sxz, syz = [np.asarray(s, dtype=np.float64) for s in S]
    n2 = sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz)
    n = np.sqrt(n2)
    pos = n > 0
    inv = np.where(pos, 1.0 / np.where(pos, n, 1.0), 0.0)
    x = np.empty((NF,) + n.shape)
    x[0] = np.where(pos, np.log10(np.where(pos, n, 1.0)), -30.0)
    for k, s in enumerate((sxx, syy, szz, sxy, sxz, syz)):
        x[1 + k] = s * inv
    tr = sxx + syy + szz
    x[7] = tr * inv
    t3 = tr * (1.0 / 3.0)
    dxx, dyy, dzz = sxx - t3, syy - t3, szz - t3
    d2 = dxx * dxx + dyy * dyy + dzz * dzz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz)
    dn = np.sqrt(d2)
    d3 = (dxx ** 3 + dyy ** 3 + dzz ** 3 + 3.0 * dxx * (sxy * sxy + sxz * sxz) + 3.0 * dyy * (sxy * sxy + syz * syz) +
          3.0 * dzz * (sxz * sxz + syz * syz) + 6.0 * sxy * sxz * syz)
    dpos = dn > 0
    x[8] = np.where(dpos, 2.449489742783178 * d3 / np.where(dpos, dn, 1.0) ** 3, 0.0)
    x[9] = 100.0 * (np.asarray(rho, dtype=np.float64) - 1.0)
    nu0 = (tau0 - 0.5) / 3.0
    x[10] = np.where(pos, np.log10(np.where(pos, n, 1.0) / nu0), -30.0)
    x[11] = dn * inv
    return x


def strain_rate(ux, uy, uz):
    """Symmetric velocity gradient of a periodic field by central differences
    (unit spacing); arrays shaped (nz, ny, nx).  Returns (xx, yy, zz, xy, xz, yz)."""
    def d(a, axis):
        return 0.5 * (np.roll(a, -1, axis) - np.roll(a, 1, axis))

    gx = [d(c, 2) for c in (ux, uy, uz)]
    gy = [d(c, 1) for c in (ux, uy, uz)]
    gz = [d(c, 0) for c in (ux, uy, uz)]
    return (gx[0], gy[1], gz[2], 0.5 * (gy[0] + gx[1]), 0.5 * (gz[0] + gx[2]), 0.5 * (gz[1] + gy[2]))


def sgs_targets(filtered):
    """(S, C*, |S|) on the coarse grid from lbm_filter output
    (rho, u, v, w, uu, vv, ww, uv, uw, vw)."""
    rho, u, v, w, uu, vv, ww, uv, uw, vw = filtered
    S = strain_rate(u, v, w)
    t = [uu - u * u, vv - v * v, ww - w * w, uv - u * v, uw - u * w, vw - v * w]
    tr = (t[0] + t[1] + t[2]) / 3.0
    t[0], t[1], t[2] = t[0] - tr, t[1] - tr, t[2] - tr
    sxx, syy, szz, sxy, sxz, syz = S
    str_ = (sxx + syy + szz) / 3.0
    dS = (sxx - str_, syy - str_, szz - str_, sxy, sxz, syz)
    tS = t[0] * dS[0] + t[1] * dS[1] + t[2] * dS[2] + 2.0 * (t[3] * dS[3] + t[4] * dS[4] + t[5] * dS[5])
    SS = dS[0] ** 2 + dS[1] ** 2 + dS[2] ** 2 + 2.0 * (dS[3] ** 2 + dS[4] ** 2 + dS[5] ** 2)
    nS = np.sqrt(sxx ** 2 + syy ** 2 + szz ** 2 + 2.0 * (sxy ** 2 + sxz ** 2 + syz ** 2))
    with np.errstate(divide="ignore", invalid="ignore"):
        nu_t = np.where(SS > 0, -tS / (2.0 * SS), 0.0)
        C = np.where(nS > 0, nu_t / nS, 0.0)
    return S, C, nS


def training_set(filtered, tau0):
    """(X [n, 12], y [n], w [n]): features, target coefficient and dissipation
    weight |S|^3 (normalised to mean 1) of every coarse cell with |S| > 0."""
    S, C, nS = sgs_targets(filtered)
    X = strain_features(S, filtered[0], tau0)
    keep = (nS > 0).ravel()
    X = X.reshape(NF, -1)[:, keep].T
    y = C.ravel()[keep]
    w = nS.ravel()[keep] ** 3
    w = w / w.mean()
    return X, y, w


def fit_normalisation(X, q=(0.0, 1.0)):
    """norm = [mu, sd, lo, hi] (4 x 12): standardisation and the training
    envelope as quantiles of the standardised features (default: the full
    range the network was trained on)."""
    mu = X.mean(axis=0)
    sd = X.std(axis=0)
    sd = np.where(sd > 1e-6, sd, 1.0)
    Z = (X - mu) / sd
    lo = np.quantile(Z, q[0], axis=0)
    hi = np.quantile(Z, q[1], axis=0)
    return np.stack([mu, sd, lo, hi]).astype(np.float32)


def normalise(X, norm):
    return ((X - norm[0]) / norm[1]).astype(np.float32)


# --- experiment -----------------------------------------------------------

def _log(msg):
    print(f"[closure] {msg}", flush=True)


def fine_run(n, k0, seed, nu, steps, snap_every, factor, u_rms=0.04, storage=abi.STORAGE_F32):
    """Decaying turbulence on n^3 (cumulant, no closure); yields
    (step, filtered fields) every snap_every steps, step 0 included."""
    tau = 3.0 * nu + 0.5
    t0 = time.time()
    ux, uy, uz = decaying_turbulence(n, k0=k0, u_rms=u_rms, seed=seed, dtype=np.float64)
    _log(f"fine {n}^3 k0={k0} seed={seed} field {time.time() - t0:.1f} s")
    flags = np.ones((n, n, n), np.int32)
    lat = Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_NONE, storage=storage,
                            arith=abi.ARITH_FAST, compat=0, periodic=(1, 1, 1), tau=tau))
    lat.init_equilibrium(None, ux, uy, uz)
    del ux, uy, uz
    done = 0
    t0 = time.time()
    while True:
        yield done, lat.filter(factor)
        if done >= steps:
            break
        k = min(snap_every, steps - done)
        lat.step(k)
        done += k
        lat.sync()
        _log(f"fine {n}^3 step {done}/{steps} ({time.time() - t0:.1f} s)")
    lat.close()


def kinetic_energy(u, v, w):
    return 0.5 * float(np.mean(u * u + v * v + w * w))


def les_run(filtered0, nu, closure, steps, every, c_les=0.17, params=None, norm=None, refresh=1):
    """Coarse LES from a filtered field: returns (steps, K(t)) and the final
    velocity, with closure 'smagorinsky', 'mlp' (network evaluated every
    `refresh` steps) or 'none' (the molecular rate alone, no sub-grid model)."""
    n = filtered0.shape[1]
    flags = np.ones((n, n, n), np.int32)
    cl = {"mlp": abi.CLOSURE_MLP, "none": abi.CLOSURE_NONE}.get(closure, abi.CLOSURE_SMAGORINSKY)
    lat = Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2015, closure=cl, storage=abi.STORAGE_F32,
                            arith=abi.ARITH_FAST, compat=0, periodic=(1, 1, 1), tau=3.0 * nu + 0.5, c_les=c_les))
    if closure == "mlp":
        lat.set_closure(params, norm)
        lat.set_closure_refresh(refresh)
    lat.init_equilibrium(filtered0[0], filtered0[1], filtered0[2], filtered0[3])
    ts, K = [], []
    done = 0
    frac = None
    while True:
        _, u, v, w = lat.macroscopic()
        ts.append(done)
        K.append(kinetic_energy(u, v, w))
        if done >= steps:
            break
        lat.step(min(every, steps - done))
        done = min(done + every, steps)
    if closure == "mlp":
        feats, env, _ = lat.closure_features()
        frac = float(env.mean())
        out_of = [float(np.mean((feats[k] < norm[2][k]) | (feats[k] > norm[3][k]))) for k in range(NF)]
        _log("outside the envelope per feature: " + " ".join(f"{v:.3f}" for v in out_of))
    t0 = time.time()
    lat.step(200)
    lat.sync()
    ms = (time.time() - t0) / 200 * 1e3
    lat.close()
    return np.array(ts), np.array(K), (u, v, w), ms, frac


def dissipation(t, K):
    """-dK/dt by central differences on the sample times."""
    t = np.asarray(t, dtype=np.float64)
    return -np.gradient(K, t)


def run_experiment(fine_n=512, train_k0=8, train_seed=20240602, eval_k0=4, eval_seed=20240601, nu_fine=4e-3,
                   train_steps=3000, train_snaps=(1000, 2000, 3000), eval_steps=4000, eval_every=100, batch=32768,
                   adam_steps=3000, lr=1e-3, c_les=0.17, seed=7, refresh=10, extra_envelopes=((0.0005, 0.9995),),
                   out_json=None, f_eval=2):
    """The learned-closure experiment: data from a fine run filtered by 4,
    training, then LES at fine_n/f_eval against the fine run filtered by f_eval."""
    res = {"fine_n": fine_n, "nu_fine": nu_fine, "c_les": c_les, "batch": batch}
    # 1. training data
    f_train = 4
    nu_c = nu_fine * f_train / (f_train * f_train)
    tau_c = 3.0 * nu_c + 0.5
    Xs, ys, ws = [], [], []
    for step, filt in fine_run(fine_n, train_k0, train_seed, nu_fine, train_steps, 500, f_train):
        if step in train_snaps:
            X, y, w = training_set(filt, tau_c)
            Xs.append(X)
            ys.append(y)
            ws.append(w)
            _log(f"train snapshot step {step}: {len(y)} samples, <C*>_w = {np.average(y, weights=w):.4f}")
    X = np.concatenate(Xs)
    y = np.concatenate(ys)
    w = np.concatenate(ws)
    norm = fit_normalisation(X)
    Xn = normalise(X, norm)
    res["train_samples"] = int(len(y))
    res["target_weighted_mean_C"] = float(np.average(y, weights=w))
    # 2. training
    m = Mlp(seed=seed)
    m.set_data(Xn, y, w)
    t0 = time.time()
    loss = m.train(adam_steps, batch, lr, c_les, seed=seed)
    res["train_seconds"] = time.time() - t0
    res["loss_first"] = float(loss[:50].mean())
    res["loss_last"] = float(loss[-50:].mean())
    # baseline loss of the constant Smagorinsky coefficient on the same weights
    res["loss_smagorinsky"] = float(np.average((c_les - y) ** 2, weights=w))
    params = m.params()
    _log(f"trained {adam_steps} steps in {res['train_seconds']:.1f} s: loss {res['loss_first']:.3e} -> "
         f"{res['loss_last']:.3e} (Smagorinsky {res['loss_smagorinsky']:.3e})")
    # 3. evaluation: fine run filtered by f_eval vs LES on the coarse grid
    res["f_eval"] = f_eval
    nu_e = nu_fine * f_eval / (f_eval * f_eval)
    fine_t, fine_K, fine_final, filt0 = [], [], None, None
    for step, filt in fine_run(fine_n, eval_k0, eval_seed, nu_fine, eval_steps, eval_every * f_eval, f_eval):
        if filt0 is None:
            filt0 = filt
        fine_t.append(step / f_eval)
        fine_K.append(kinetic_energy(filt[1], filt[2], filt[3]))
        fine_final = (filt[1], filt[2], filt[3])
    fine_t = np.array(fine_t)
    fine_K = np.array(fine_K)
    les_steps = eval_steps // f_eval
    out = {}
    # controls: the reference's constant, the constant fitted to the training
    # targets (their dissipation-weighted mean), and no sub-grid model at all
    c_fit = res["target_weighted_mean_C"]
    runs = (("smagorinsky", "smagorinsky", 1, c_les), ("smagorinsky_fitted", "smagorinsky", 1, c_fit),
            ("no_model", "none", 1, c_les), ("mlp", "mlp", 1, c_les),
            (f"mlp_refresh{refresh}", "mlp", refresh, c_les))
    for name, kind, every, c_run in runs:
        t, K, fin, ms, frac = les_run(filt0, nu_e, kind, les_steps, eval_every, c_run, params, norm, every)
        eps = dissipation(t, K)
        eps_ref = dissipation(fine_t, fine_K)
        # compare after the initial transient (first tenth of the run)
        sel = t >= 0.1 * les_steps
        rel = float(np.mean(np.abs(eps[sel] - eps_ref[sel]) / np.abs(eps_ref[sel])))
        kf, Ef = energy_spectrum(*fine_final)
        kl, El = energy_spectrum(*fin)
        kmax = len(Ef) // 2
        spec = float(np.sqrt(np.mean((np.log(El[1:kmax]) - np.log(Ef[1:kmax])) ** 2)))
        out[name] = {"dissipation_rel_err": rel, "spectrum_log_rms": spec, "ms_per_step": ms,
                     "K_final": float(K[-1]), "in_envelope": frac, "c_les": c_run if kind != "none" else None}
        _log(f"{name}: dissipation rel err {rel:.3f}, spectrum log-rms {spec:.3f}, {ms:.3f} ms/step"
             + (f", envelope {frac:.3f}" if frac is not None else ""))
    # the same network with wider training envelopes (fewer Smagorinsky fallbacks)
    for q in extra_envelopes:
        nq = fit_normalisation(X, q=q)
        t, K, fin, ms, frac = les_run(filt0, nu_e, "mlp", les_steps, eval_every, c_les, params, nq, refresh)
        eps = dissipation(t, K)
        eps_ref = dissipation(fine_t, fine_K)
        sel = t >= 0.1 * les_steps
        rel = float(np.mean(np.abs(eps[sel] - eps_ref[sel]) / np.abs(eps_ref[sel])))
        kl, El = energy_spectrum(*fin)
        spec = float(np.sqrt(np.mean((np.log(El[1:kmax]) - np.log(Ef[1:kmax])) ** 2)))
        key = f"mlp_refresh{refresh}_envelope_q{q[0]:g}"
        out[key] = {"dissipation_rel_err": rel, "spectrum_log_rms": spec, "ms_per_step": ms, "K_final": float(K[-1]),
                    "in_envelope": frac, "envelope_quantiles": list(q)}
        _log(f"{key}: dissipation rel err {rel:.3f}, spectrum log-rms {spec:.3f}, envelope {frac:.3f}")
    out["fine_K_final"] = float(fine_K[-1])
    res["eval"] = out
    res["mlups_cost"] = out["mlp"]["ms_per_step"] / out["smagorinsky"]["ms_per_step"] - 1.0
    res["refresh"] = refresh
    res["mlups_cost_refresh"] = out[f"mlp_refresh{refresh}"]["ms_per_step"] / out["smagorinsky"]["ms_per_step"] - 1.0
    if out_json:
        with open(out_json, "w") as fh:
            json.dump(res, fh, indent=1)
        np.save(out_json.replace(".json", "_params.npy"), params)
        np.save(out_json.replace(".json", "_norm.npy"), norm)
    return res


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="learned-closure experiment")
    ap.add_argument("--fine", type=int, default=512)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--eval-steps", type=int, default=4000)
    ap.add_argument("--adam-steps", type=int, default=3000)
    ap.add_argument("--refresh", type=int, default=10)
    ap.add_argument("--f-eval", type=int, default=2, help="filter width of the evaluation (LES at fine/f)")
    ap.add_argument("--eval-every", type=int, default=100)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    snaps = tuple(int(a.train_steps * k / 3) for k in (1, 2, 3))
    r = run_experiment(fine_n=a.fine, train_steps=a.train_steps, train_snaps=snaps, eval_steps=a.eval_steps,
                       eval_every=a.eval_every, adam_steps=a.adam_steps, refresh=a.refresh, out_json=a.out,
                       f_eval=a.f_eval)
    print(json.dumps(r))
