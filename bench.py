#!/usr/bin/env python3
"""Headline benchmark: D3Q27 cumulant + Smagorinsky LES, 256^3, FP32, one MI355X.

Workload (config C2): periodic decaying turbulence on 256^3 cells, initial
velocity from seeded random Fourier modes with E(k) ~ k^4 exp(-2 (k/k0)^2),
k0 = 4, rms u = 0.04 (seed 20240601), rho = 1, f = f_eq; cumulant-2015
collision with the Smagorinsky rate (C = 0.17 in the reference formula),
tau0 = 0.5015 (nu0 = 5e-4).  A step is one fused collide-stream launch over
all 16,777,216 cells; populations stay resident in HBM.

Prints ONE JSON line.  `value` is MLUPS (million cell updates per second)
from the wall clock around K steps bracketed by device synchronisation; the
roofline block uses the HIP-event time of the collide-stream kernel on the
stream it is launched on.  The CPU baseline is the CPU oracle (a restatement
of the reference's unfused two-lattice step) on a bounded sample of the same
workload, on this host's cores.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from lb_amd import Lattice, abi, make_desc  # noqa: E402
from lb_amd.fields import decaying_turbulence  # noqa: E402

METRIC = "MLUPS + % HBM roofline, D3Q27 cumulant 256³; NN closure vs Smagorinsky err"
HBM_PEAK_GBS = 8000.0
SEED = 20240601
TAU0 = 0.5015          # nu0 = 5e-4
C_LES = 0.17
STORAGES = {v: k for k, v in abi.STORAGE_NAMES.items()}


def profiled_traffic(storage_name, n):
    """HBM bytes per launch of this kernel/workload from the committed
    rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel
    of known size), or None when that workload has not been profiled."""
    path = os.path.join(ROOT, "profiles", f"k_step_{storage_name}_c2_{n}cube.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        d = json.load(fh)
    return d["traffic_bytes_per_launch"]["mean"], os.path.relpath(path, ROOT)


def closure_result():
    """Dissipation and spectrum errors of the learned closure and its
    controls against the filtered fine run, quoted from the committed
    experiment summary (the experiment itself takes minutes)."""
    path = os.path.join(ROOT, "profiles", "closure_c5_512to256.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    runs = {k: v for k, v in d["eval"].items() if isinstance(v, dict)}
    return {"source": os.path.relpath(path, ROOT),
            "dissipation_rel_err": {k: round(v["dissipation_rel_err"], 4) for k, v in runs.items()},
            "spectrum_log_rms": {k: round(v["spectrum_log_rms"], 3) for k, v in runs.items()},
            "step_cost_vs_smagorinsky": {k: round(v["ms_per_step"] / runs["smagorinsky"]["ms_per_step"] - 1, 4)
                                         for k, v in runs.items() if k != "smagorinsky"}}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


C1_N, C1_STEPS, C1_TAU, C1_LID = 64, 2000, 0.524, 0.05


def c1_cavity(n=C1_N):
    """Lid-driven cavity of the C1 workload, meshed by the host mesher: n^3
    fluid cells inside a one-cell solid halo, the top fluid layer (y = n) a
    moving-wall inlet at (0.05, 0, 0) with outward normal +y, half-way
    bounce-back on the other walls.  Returns (flags, bc, initial u_x)."""
    from lb_amd import host

    g = np.zeros((n + 2, n + 2, n + 2), np.int32)
    g[1:-1, 1:-1, 1:-1] = 1
    m = host.Mesher(g, 1.0, 1.0, (0.0, 0.0, 0.0))
    m.wall()
    # cell centres at j + 0.5 - 1: the slab [n - 0.75, n - 0.25] holds j = n only
    m.inlet((-1.0, n - 0.75, -1.0), (n + 1.0, n - 0.25, n + 1.0), (-1.0, n - 0.75, n + 1.0), -1, (0.0, 1.0, 0.0),
            (C1_LID, 0.0, 0.0), abi.BC_INLET_MOVING_WALL)
    flags, bc = m.compile()
    ux = np.zeros(flags.shape)
    for b in bc:
        if b.kind == abi.BC_INLET_MOVING_WALL:
            ux[b.z, b.y, b.x] = b.ux
    return flags, bc, ux


def c1_desc(flags, bc):
    return make_desc(flags, bc, collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_NONE, storage=abi.STORAGE_F64,
                     arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE, tau=C1_TAU)


def c1_gpu(flags, bc, ux):
    """The C1 cavity on the GPU: FP64 storage, literal arithmetic, all steps."""
    L = Lattice(c1_desc(flags, bc))
    L.init_equilibrium(None, ux, None, None)
    L.step(10)
    L.sync()
    L.init_equilibrium(None, ux, None, None)
    t0 = time.perf_counter()
    L.step(C1_STEPS)
    L.sync()
    wall = time.perf_counter() - t0
    kms = L.last_kernel_ms()
    m = L.macroscopic()
    L.close()
    return wall, kms, m


def c1_cpu(flags, bc, ux):
    """The C1 cavity on the CPU oracle (reference order, FP64), all steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import OracleLattice, oracle_lib

    o = OracleLattice(c1_desc(flags, bc))
    o.init_equilibrium(None, ux, None, None)
    t0 = time.perf_counter()
    o.step(C1_STEPS)
    dt = time.perf_counter() - t0
    m = o.macroscopic()
    o.close()
    return dt, oracle_lib().orc_get_threads(), m


def make_lattice(n, storage, fields, arith=abi.ARITH_FAST, collision=abi.COLLISION_CUM2015):
    flags = np.ones((n, n, n), np.int32)
    d = make_desc(flags, [], collision=collision, closure=abi.CLOSURE_SMAGORINSKY, storage=storage,
                  arith=arith, compat=abi.COMPAT_REFERENCE, periodic=(1, 1, 1), tau=TAU0, c_les=C_LES)
    L = Lattice(d)
    L.init_equilibrium(None, *fields)
    return L


def closure_network(seed=7):
    """A seeded closure network (untrained weights, non-zero output layer) and
    an open envelope: the step cost does not depend on the weights."""
    rng = np.random.default_rng(seed)
    p = np.zeros(abi.MLP_PARAMS, np.float32)
    p[:768] = rng.standard_normal(768) / np.sqrt(12.0)
    p[832:4928] = rng.standard_normal(4096) / 8.0
    p[4992:5056] = 0.05 * rng.standard_normal(64)
    norm = np.stack([np.zeros(12), np.ones(12), np.full(12, -1e30), np.full(12, 1e30)]).astype(np.float32)
    return p, norm


def run_gpu(n, storage, fields, steps, warmup, arith=abi.ARITH_FAST, mlp=False, refresh=1,
            collision=abi.COLLISION_CUM2015):
    if mlp:
        flags = np.ones((n, n, n), np.int32)
        L = Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_MLP, storage=storage,
                              arith=abi.ARITH_FAST, compat=abi.COMPAT_REFERENCE, periodic=(1, 1, 1), tau=TAU0,
                              c_les=C_LES))
        L.set_closure(*closure_network())
        L.set_closure_refresh(refresh)
        L.init_equilibrium(None, *fields)
    else:
        L = make_lattice(n, storage, fields, arith, collision)
    L.step(warmup)
    L.sync()
    t0 = time.perf_counter()
    L.step(steps)
    L.sync()
    wall = time.perf_counter() - t0
    kms = L.last_kernel_ms()
    mass = L.conservation()
    L.close()
    cells = n ** 3
    return {"wall_s": wall, "ms_per_step": wall / steps * 1e3, "kernel_ms": kms,
            "mlups": cells * steps / wall / 1e6, "mass": float(mass[0]), "momentum": [float(m) for m in mass[1:]]}


def cpu_baseline(n, fields, steps, warmup):
    """CPU oracle (reference-order AB step, FP32) on the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import OracleLattice, oracle_lib

    flags = np.ones((n, n, n), np.int32)
    d = make_desc(flags, [], collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY,
                  storage=abi.STORAGE_F32, arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE,
                  periodic=(1, 1, 1), tau=TAU0, c_les=C_LES)
    o = OracleLattice(d)
    o.init_equilibrium(None, *fields)
    o.step(warmup)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    cores = oracle_lib().orc_get_threads()
    o.close()
    return {"value": round(n ** 3 * steps / dt / 1e6, 3), "unit": "MLUPS", "cores": cores, "kind": "port",
            "sample": f"CPU oracle (reference unfused AB order: omega, collision, streaming, moments), FP32, "
                      f"same {n}^3 workload, {warmup} warm-up + {steps} timed steps, {cores} OpenMP threads, "
                      f"{dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--storage", default="f32", choices=sorted(STORAGES))
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true", help="skip the other storage formats")
    ap.add_argument("--no-c1", action="store_true", help="skip the C1 cavity row")
    ap.add_argument("--closure-only", action="store_true",
                    help="time only the plain FP32 step and the learned-closure step (network every step), "
                         "print one line each and stop (for profiling)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the collide-stream core runs on a single MI355X; "
              "multi-GPU slab decomposition is a design, not an implementation) -- run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    n_dev, dev = abi.device_info()
    if n_dev <= 0:
        raise SystemExit(f"no HIP device visible: {dev}")
    log(f"device: {dev}")
    probe_gbs = 0.0
    g = abi.load_core()
    gb, ms = ctypes.c_double(), ctypes.c_double()
    if g.lbm_bandwidth_probe(1 << 30, 20, ctypes.byref(gb), ctypes.byref(ms)) == abi.LBM_OK:
        probe_gbs = gb.value
    log(f"copy-kernel bandwidth: {probe_gbs:.0f} GB/s")

    n = args.n
    t0 = time.perf_counter()
    fields = decaying_turbulence(n, k0=4.0, u_rms=0.04, seed=SEED)
    log(f"initial field {n}^3 generated in {time.perf_counter() - t0:.1f} s")

    storage = STORAGES[args.storage]
    sbytes = abi.STORAGE_BYTES[storage]
    res = run_gpu(n, storage, fields, args.steps, args.warmup)
    log(f"{args.storage}: {res['mlups']:.0f} MLUPS, kernel {res['kernel_ms']:.4f} ms/step, "
        f"wall {res['ms_per_step']:.4f} ms/step, mass {res['mass']:.6e}")
    if not np.isfinite(res["mass"]):
        raise SystemExit("non-finite mass after the timed run")
    if args.closure_only:
        r = run_gpu(n, abi.STORAGE_F32, fields, args.steps, args.warmup, mlp=True, refresh=1)
        log(f"f32 + learned closure (every 1): kernel {r['kernel_ms']:.4f} ms/step "
            f"({r['kernel_ms'] / res['kernel_ms'] - 1:+.1%} over the plain step)")
        return

    cells = n ** 3
    bytes_per_launch = cells * 2 * 27 * sbytes
    achieved = bytes_per_launch / (res["kernel_ms"] * 1e-3) / 1e9
    traffic, traffic_src = profiled_traffic(args.storage, n)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_quoted_from": traffic_src,
                "algorithmic_bytes_per_launch": bytes_per_launch,
                "kernel": "k_step (fused collide-stream, A-A)", "bytes_per_cell": 2 * 27 * sbytes,
                "kernel_ms": round(res["kernel_ms"], 5), "copy_kernel_gbs": round(probe_gbs, 1)}

    variants = {}
    if not args.no_variants:
        vsteps = max(200, args.steps // 5)
        for name in ("f16dev", "f16dev_sr", "bf16dev", "bf16dev_sr", "f64"):
            if name == args.storage:
                continue
            r = run_gpu(n, STORAGES[name], fields, vsteps, 20)
            sb = abi.STORAGE_BYTES[STORAGES[name]]
            ach = cells * 54 * sb / (r["kernel_ms"] * 1e-3) / 1e9
            variants[name] = {"mlups": round(r["mlups"], 1), "kernel_ms": round(r["kernel_ms"], 5),
                              "achieved_gbs": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
                              "steps": vsteps, "bytes_per_cell": 54 * sb,
                              "vs_f32_mlups": round(r["mlups"] / res["mlups"], 3)}
            log(f"{name}: {r['mlups']:.0f} MLUPS, {ach:.0f} GB/s")
        # the Geier-2017 operator (limited third-order rates, fourth-order
        # A/B terms) on the same workload
        for name in ("f32", "f16dev_sr"):
            r = run_gpu(n, STORAGES[name], fields, vsteps, 20, collision=abi.COLLISION_CUM2017)
            sb = abi.STORAGE_BYTES[STORAGES[name]]
            ach = cells * 54 * sb / (r["kernel_ms"] * 1e-3) / 1e9
            variants[f"{name}_cum2017"] = {"mlups": round(r["mlups"], 1), "kernel_ms": round(r["kernel_ms"], 5),
                                           "achieved_gbs": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
                                           "steps": vsteps, "vs_f32_mlups": round(r["mlups"] / res["mlups"], 3)}
            log(f"{name} cum2017: {r['mlups']:.0f} MLUPS, {ach:.0f} GB/s")
        # learned closure fused into the same step (FP32 storage), network
        # evaluated every step and every 10 steps (coefficient held between)
        for every in (1, 10):
            r = run_gpu(n, abi.STORAGE_F32, fields, vsteps, 20, mlp=True, refresh=every)
            key = "f32_mlp_closure" if every == 1 else f"f32_mlp_closure_refresh{every}"
            variants[key] = {"mlups": round(r["mlups"], 1), "kernel_ms": round(r["kernel_ms"], 5), "steps": vsteps,
                             "refresh": every,
                             "cost_vs_smagorinsky": round(r["kernel_ms"] / res["kernel_ms"] - 1, 4)}
            log(f"f32 + learned closure (every {every}): {r['mlups']:.0f} MLUPS "
                f"({variants[key]['cost_vs_smagorinsky']:+.1%})")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(n, fields, args.cpu_steps, 2)
        log(f"cpu baseline: {cpu['value']} MLUPS on {cpu['cores']} threads")

    # C1: the lid-driven cavity, FP64 literal on the GPU and on the CPU oracle
    c1 = None
    if not args.no_c1:
        flags1, bc1, ux1 = c1_cavity()
        cells1 = int((flags1 != 0).sum())
        wall1, kms1, mg = c1_gpu(flags1, bc1, ux1)
        c1 = {"workload": f"C1 lid-driven cavity {C1_N}^3 (+1 solid halo), Re=400, lid u={C1_LID}, tau={C1_TAU}, "
                          f"cumulant-2015 laminar, FP64 literal, {C1_STEPS} steps",
              "cells": cells1, "gpu_mlups": round(cells1 * C1_STEPS / wall1 / 1e6, 1),
              "gpu_kernel_ms": round(kms1, 5)}
        if not args.no_cpu_baseline:
            dt1, cores1, mo = c1_cpu(flags1, bc1, ux1)
            fl = flags1 != 0
            num = np.sqrt(sum(np.sum((a[fl] - b[fl]) ** 2) for a, b in zip(mg[1:], mo[1:])))
            den = np.sqrt(sum(np.sum(b[fl] ** 2) for b in mo[1:]))
            c1.update({"cpu_mlups": round(cells1 * C1_STEPS / dt1 / 1e6, 2), "cpu_cores": cores1,
                       "cpu_kind": "port", "gpu_vs_cpu_rel_l2_u": float(num / den)})
        log(f"C1 cavity: GPU {c1['gpu_mlups']} MLUPS" +
            (f", CPU {c1['cpu_mlups']} MLUPS on {c1['cpu_cores']} threads, rel-L2(u) "
             f"{c1['gpu_vs_cpu_rel_l2_u']:.1e}" if "cpu_mlups" in c1 else ""))

    out = {
        "metric": METRIC,
        "value": round(res["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if storage == abi.STORAGE_F32 else ("f64" if storage == abi.STORAGE_F64 else "f32"),
        "data": "synthetic",
        "config": {"workload": f"C2 decaying turbulence {n}^3 periodic, D3Q27 cumulant-2015 + Smagorinsky "
                               f"(C=0.17), {args.storage} population storage, fused A-A collide-stream",
                   "cells": cells, "storage": args.storage, "tau0": TAU0, "c_les": C_LES,
                   "spectrum": "E(k)~k^4 exp(-2(k/4)^2), u_rms=0.04", "seed": SEED, "device": dev},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "variants": variants,
        "c1": c1,
        "quoted_from_profiles": {"closure_c5": closure_result()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
