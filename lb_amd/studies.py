"""Measured studies on the GPU (host orchestration over the C ABI).

c3  Taylor-Green vortex 384^3, Re 1600, cumulant: FP64 storage as the
    reference run against FP32, FP16-deviation and BF16-deviation storage,
    each 16-bit format with round-to-nearest and with stochastic rounding
    (all FP32 arithmetic except FP64); rel-L2 of u and rho and the
    kinetic-energy error over 10 000 steps, MLUPS per storage.
c4  flow past a voxelized obstacle, 512 x 256 x 256: synthetic OFF scene,
    the host voxelizer and boundary mesher, equilibrium velocity inlet at
    x = 0, interpolated density outlet at x = 511, index wrap along y and z,
    cumulant + Smagorinsky; drag from the wall-link momentum exchange, FP32
    and FP16-deviation storage (stochastic and nearest rounding) against FP64
    storage, MLUPS.
snapshot  output overlap of the host driver: a lid-driven cavity (cumulant +
    Smagorinsky, FP32) run with VTK output every `every` steps, with the
    synchronous read-back and with the asynchronous snapshot path
    (async_output), against the same run without output; wall times, the
    share of the stepping time that the asynchronous path hides behind the
    file writes, and a check that both write identical files.

usage: python -m lb_amd.studies c3|c4|snapshot [--steps N] [--out file.json]
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
import tempfile
import time

import numpy as np

from . import abi, host
from .core import Lattice, make_desc
from .fields import taylor_green
from .geometry import obstacle_scene, write_off


def _log(msg):
    print(f"[study] {msg}", flush=True)


def _rel(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2)) / np.sqrt(np.sum(b * b)))


def c3(n=384, steps=10000, every=1000, u0=0.05,
       storages=("f64", "f32", "f16dev", "bf16dev", "f16dev_sr", "bf16dev_sr")):
    L = n / (2.0 * np.pi)
    nu = u0 * L / 1600.0
    tau = 3.0 * nu + 0.5
    rho0, ux, uy, uz = taylor_green(n, u0)
    flags = np.ones((n, n, n), np.int32)
    codes = {v: k for k, v in abi.STORAGE_NAMES.items()}
    lats = {}
    for s in storages:
        lats[s] = Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2015, storage=codes[s], arith=abi.ARITH_FAST,
                                    compat=0, periodic=(1, 1, 1), tau=tau))
        lats[s].init_equilibrium(rho0, ux, uy, uz)
    del rho0, ux, uy, uz
    res = {"workload": f"Taylor-Green {n}^3, u0={u0}, Re=1600, nu={nu:.6g}, tau={tau:.6g}, cumulant-2015 fast",
           "steps": steps, "records": []}
    ms = {s: [] for s in storages}
    done = 0
    while done < steps:
        k = min(every, steps - done)
        for s in storages:
            lats[s].step(k)
            lats[s].sync()
            ms[s].append(lats[s].last_kernel_ms())
        done += k
        ref = lats["f64"].macroscopic()
        Kref = 0.5 * float(np.mean(ref[1] ** 2 + ref[2] ** 2 + ref[3] ** 2))
        rec = {"step": done, "K_f64": Kref}
        for s in storages:
            if s == "f64":
                continue
            m = lats[s].macroscopic()
            K = 0.5 * float(np.mean(m[1] ** 2 + m[2] ** 2 + m[3] ** 2))
            rec[s] = {"rel_l2_u": _rel(np.stack(m[1:]), np.stack(ref[1:])), "rel_l2_rho": _rel(m[0], ref[0]),
                      "rel_l2_rho_minus_1": _rel(m[0] - 1.0, ref[0] - 1.0), "K_rel_err": abs(K - Kref) / Kref}
            del m
        res["records"].append(rec)
        _log(f"step {done}: " + ", ".join(f"{s} u {rec[s]['rel_l2_u']:.2e}" for s in storages if s != "f64"))
        del ref
    res["mlups"] = {s: n ** 3 / (float(np.median(ms[s])) * 1e-3) / 1e6 for s in storages}
    res["ms_per_step"] = {s: float(np.median(ms[s])) for s in storages}
    for s in storages:
        lats[s].close()
    return res


def c4_geometry(workdir, res=1.0, add=0.0, scene=None, u_in=0.04):
    """Voxelized flag grid and boundary list of the obstacle configuration
    (scene: (vertices, triangles, objects) of another obstacle_scene)."""
    v, f, objs = scene if scene is not None else obstacle_scene()
    path = os.path.join(workdir, "obstacle.off")
    write_off(path, v, f)
    t0 = time.time()
    grid, mn, mx = host.voxelize_off(path, res, add, (-3.1416, -2.7183, -1.4142))
    tv = time.time() - t0
    # the same inside tests on the GPU: same grid
    t0 = time.time()
    grid_gpu, _, _ = host.voxelize_off(path, res, add, (-3.1416, -2.7183, -1.4142), device=True)
    tv_gpu = time.time() - t0
    if not np.array_equal(grid, grid_gpu):
        raise RuntimeError("GPU voxelization differs from the host's")
    nz, ny, nx = grid.shape
    m = host.Mesher(grid, res, add, mn)
    m.wall()
    big = 1e4
    # cell centres sit at i + 0.5: the planes x = 0 and x = nx - 1
    m.inlet((0.2, -big, -big), (0.8, big, big), (0.8, big, -big), -1, (-1.0, 0.0, 0.0), (u_in, 0.0, 0.0),
            abi.BC_INLET_EQUILIBRIUM)
    m.outlet((nx - 0.8, -big, -big), (nx - 0.2, big, big), (nx - 0.2, big, -big), -2, (1.0, 0.0, 0.0), 1.0)
    flags, bc = m.compile()
    # analytic check of the voxelization away from the surfaces
    z, y, x = [a + 0.5 for a in np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij", sparse=True)]
    inside_any = np.zeros(grid.shape, bool)
    near = np.zeros(grid.shape, bool)
    for _, c, r in objs:
        s = np.sqrt(((x - c[0]) / r[0]) ** 2 + ((y - c[1]) / r[1]) ** 2 + ((z - c[2]) / r[2]) ** 2)
        inside_any |= s < 1
        near |= np.abs(s - 1) < 2.0 / r.min()
    far = ~near
    mism = int(np.count_nonzero(grid[far].astype(bool) == inside_any[far]))
    info = {"dims": [nx, ny, nz], "triangles": int(len(f)), "voxelize_s": tv, "voxelize_gpu_s": tv_gpu,
            "fluid_cells": int(grid.sum()),
            "solid_cells": int(grid.size - grid.sum()), "analytic_mismatch_far": mism,
            "near_surface_cells": int(near.sum()), "wall_nodes": int(sum(1 for b in bc if b.kind == abi.BC_WALL)),
            "inlet_nodes": int(sum(1 for b in bc if b.kind == abi.BC_INLET_EQUILIBRIUM)),
            "outlet_nodes": int(sum(1 for b in bc if b.kind == abi.BC_OUTLET_INTERPOLATED)),
            "obstacles": [[k, [float(t) for t in c], [float(t) for t in r]] for k, c, r in objs]}
    return flags, bc, info


def c4(steps=5000, every=250, fp64_steps=None):
    D = 48.0
    u_in = 0.04
    nu = u_in * D / 1000.0
    tau = 3.0 * nu + 0.5
    with tempfile.TemporaryDirectory() as td:
        flags, bc, info = c4_geometry(td)
    _log(f"geometry {info['dims']} voxelized in {info['voxelize_s']:.1f} s, far-field mismatches "
         f"{info['analytic_mismatch_far']}, walls {info['wall_nodes']}")
    res = {"workload": f"obstacle 512x256x256, inlet u={u_in}, outlet rho=1, y/z wrap, Re_D=1000 nu={nu:.4g}, "
                       "cumulant-2015 + Smagorinsky C=0.17", "geometry": info}
    fp64_steps = fp64_steps or steps
    ux0 = np.where(flags == 0, 0.0, u_in)
    runs = (("f32", abi.STORAGE_F32, steps), ("f16dev_sr", abi.STORAGE_F16DEV_SR, steps),
            ("f16dev", abi.STORAGE_F16DEV, steps), ("f64", abi.STORAGE_F64, fp64_steps))
    for name, st, nst in runs:
        lat = Lattice(make_desc(flags, bc, collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY,
                                storage=st, arith=abi.ARITH_FAST, compat=abi.COMPAT_REFERENCE, periodic=(0, 1, 1),
                                tau=tau, nu=nu, c_les=0.17))
        lat.init_equilibrium(None, ux0, None, None)
        t, F, ms = [], [], []
        done = 0
        t0 = time.time()
        while done < nst:
            k = min(every - 1, nst - done - 1)
            if k > 0:
                lat.step(k)
                lat.sync()
                ms.append(lat.last_kernel_ms())
            _, _, force = lat.step_diagnostics()
            done += k + 1
            t.append(done)
            F.append(force.tolist())
        _log(f"{name}: {nst} steps in {time.time() - t0:.1f} s, drag {F[-1][0]:.4e}")
        res[name] = {"steps": t, "force": F, "ms_per_step": float(np.median(ms)),
                     "mlups": float(info["fluid_cells"] / (np.median(ms) * 1e-3) / 1e6),
                     "mlups_all_cells": float(flags.size / (np.median(ms) * 1e-3) / 1e6)}
        lat.close()
    for name, _, _ in runs[:-1]:
        n = min(len(res[name]["force"]), len(res["f64"]["force"]))
        fx = np.array(res[name]["force"][:n])
        f64 = np.array(res["f64"]["force"][:n])
        half = n // 2
        res[f"drag_rel_diff_{name}_vs_f64"] = float(np.mean(np.abs(fx[half:, 0] - f64[half:, 0])) /
                                                   np.mean(np.abs(f64[half:, 0])))
    f64 = np.array(res["f64"]["force"])
    half = len(f64) // 2
    res["drag_coefficient_f64"] = float(np.mean(f64[half:, 0]) / (0.5 * u_in ** 2 * np.pi * (D / 2) ** 2))
    return res


def _cavity_case(Lh, n, steps, results_dir, **extra):
    """The host driver's lid-driven cavity (n^3 fluid cells, moving-wall lid),
    as the reference's case files set one up."""
    h = ctypes.c_void_p()
    if Lh.lbmh_case_create(ctypes.byref(h)) != 0:
        raise RuntimeError("lbmh_case_create failed")
    grid = np.zeros((n + 2,) * 3, np.int32)
    grid[1:-1, 1:-1, 1:-1] = 1
    dims = (ctypes.c_int32 * 3)(n + 2, n + 2, n + 2)
    Lh.lbmh_case_set_grid(h, dims, grid.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 1.0, 1.0,
                          (ctypes.c_float * 3)(0, 0, 0), (ctypes.c_float * 3)(n, n, n))
    tau = 0.56
    params = dict(tau=tau, ny=(tau - 0.5) / 3, rho_fyz=1.0, Cm=1.0, Cl=1.0, Cu_inverse=1.0, iterations=steps,
                  err_every_it=-1, plot_every_it=-1, iterationsMomentAvg=0, collision=abi.COLLISION_CUM2015,
                  closure=abi.CLOSURE_SMAGORINSKY, CLES=0.17, storage=abi.STORAGE_F32, arith=abi.ARITH_FAST,
                  compat=abi.COMPAT_REFERENCE)
    params.update(extra)
    for k, v in params.items():
        if Lh.lbmh_case_set(h, k.encode(), float(v)) != 0:
            raise RuntimeError(f"lbmh_case_set {k}")
    os.makedirs(results_dir, exist_ok=True)
    Lh.lbmh_case_set_results_dir(h, (str(results_dir) + "/").encode())
    ok = Lh.lbmh_case_mesh_wall(h) == 0
    lid = (ctypes.c_float * 3)(0.05, 0.0, 0.0)
    ok = ok and Lh.lbmh_case_inlet(h, (ctypes.c_double * 3)(-1.0, n - 0.75, -1.0),
                                   (ctypes.c_double * 3)(n + 1.0, n - 0.25, n + 1.0),
                                   (ctypes.c_double * 3)(-1.0, n - 0.75, n + 1.0), -1,
                                   (ctypes.c_float * 3)(0, 1, 0), lid, abi.BC_INLET_MOVING_WALL) == 0
    ok = ok and Lh.lbmh_case_compile(h) == 0 and Lh.lbmh_case_initialize(h) == 0
    if not ok:
        raise RuntimeError("cavity case set-up failed")
    return h


def snapshot(n=128, every=2000, plots=5):
    Lh = host.load_host()
    steps = every * plots
    res = {"workload": f"lid-driven cavity {n}^3 (moving-wall lid u=0.05), cumulant-2015 + Smagorinsky, FP32, "
                       f"{steps} steps, VTK 'variables' output every {every} steps", "cells": n ** 3}
    digests = {}
    with tempfile.TemporaryDirectory() as td:
        for name, extra in (("no_output", {}), ("sync", dict(plot_every_it=every, async_output=0)),
                            ("async", dict(plot_every_it=every, async_output=1))):
            d = os.path.join(td, name)
            h = _cavity_case(Lh, n, steps, d, **extra)
            t0 = time.perf_counter()
            rc = Lh.lbmh_case_run(h)
            wall = time.perf_counter() - t0
            Lh.lbmh_case_destroy(h)
            if rc != 0:
                raise RuntimeError(f"{name}: lbmh_case_run returned {rc}")
            files = sorted(f for f in os.listdir(d) if f.startswith("variables."))
            digests[name] = {f: hashlib.sha256(open(os.path.join(d, f), "rb").read()).hexdigest() for f in files}
            size = sum(os.path.getsize(os.path.join(d, f)) for f in files)
            res[name] = {"wall_s": wall, "files": len(files), "bytes_written": size}
            _log(f"{name}: {wall:.2f} s, {len(files)} files, {size / 1e6:.0f} MB")
            for f in files:
                os.remove(os.path.join(d, f))
    res["files_identical"] = digests["sync"] == digests["async"] and len(digests["sync"]) > 0
    t0, ts, ta = res["no_output"]["wall_s"], res["sync"]["wall_s"], res["async"]["wall_s"]
    res["stepping_hidden_fraction"] = (ts - ta) / t0
    _log(f"files identical: {res['files_identical']}; stepping time hidden behind output: "
         f"{res['stepping_hidden_fraction']:.0%}")
    return res


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("study", choices=["c3", "c4", "snapshot"])
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--n", type=int, default=384)
    ap.add_argument("--every", type=int, default=1000)
    ap.add_argument("--storages", default=None, help="comma-separated, c3 only (default: all)")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    if a.study == "c3":
        kw = {"storages": tuple(a.storages.split(","))} if a.storages else {}
        r = c3(n=a.n, steps=a.steps or 10000, every=a.every, **kw)
    elif a.study == "snapshot":
        r = snapshot(n=a.n if a.n != 384 else 128, every=a.every if a.every != 1000 else 2000)
    else:
        r = c4(steps=a.steps or 5000)
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(r, fh, indent=1)
    print(json.dumps({k: v for k, v in r.items() if k not in ("records",)})[:4000])
