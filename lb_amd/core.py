"""Python handle on one lattice context of the HIP core (C ABI: include/lbm_core.h)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import abi


class LbmError(RuntimeError):
    def __init__(self, status: int, where: str):
        lib = abi.load_core()
        msg = lib.lbm_status_str(status).decode()
        super().__init__(f"{where}: {abi.STATUS_NAMES.get(status, status)} ({msg})")
        self.status = status


def _check(status: int, where: str):
    if status != abi.LBM_OK:
        raise LbmError(status, where)


def make_desc(flags: np.ndarray, bc=None, *, collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_NONE,
              storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE,
              periodic=(0, 0, 0), tau=0.6, nu=0.0, c_les=0.0, rho0=1.0, sponge=None, wall_mesh=False):
    """Build an lbm_desc.  flags: int32 array shaped (nz, ny, nx) (x fastest in memory)."""
    flags = np.ascontiguousarray(flags, dtype=np.int32)
    nz, ny, nx = flags.shape
    d = abi.Desc()
    d.nx, d.ny, d.nz = nx, ny, nz
    d.periodic[:] = [int(bool(p)) for p in periodic]
    d.collision, d.closure, d.storage, d.arith, d.compat = collision, closure, storage, arith, compat
    d.tau, d.nu, d.c_les, d.rho0 = float(tau), float(nu), float(c_les), float(rho0)
    d.wall_mesh = int(bool(wall_mesh))   # BounceBackWallHalfMesh: flag-2 cells instead of the wall list
    if sponge is not None:   # (omegaDumpingLow, omegaDumpingHigh) of the outlet omega damping
        d.outlet_sponge, d.sponge_low, d.sponge_high = 1, float(sponge[0]), float(sponge[1])
    d.flags = flags.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    bc = bc if bc is not None else []
    if len(bc):
        arr = (abi.BcNode * len(bc))(*bc)
        d.bc = ctypes.cast(arr, ctypes.POINTER(abi.BcNode))
    else:
        arr = None
        d.bc = None
    d.n_bc = len(bc)
    # keep the backing buffers alive with the descriptor
    d._keep = (flags, arr)
    return d


class Lattice:
    """One D3Q27 lattice resident in HBM, advanced by the fused collide-stream kernel."""

    def __init__(self, desc: abi.Desc):
        self._lib = abi.load_core()
        self.desc = desc
        self.shape = (desc.nz, desc.ny, desc.nx)
        self.n = desc.nx * desc.ny * desc.nz
        h = ctypes.c_void_p()
        _check(self._lib.lbm_create(ctypes.byref(desc), ctypes.byref(h)), "lbm_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state ------------------------------------------------------------
    def init_equilibrium(self, rho=None, ux=None, uy=None, uz=None):
        arrs = [None if a is None else np.ascontiguousarray(np.broadcast_to(a, self.shape), dtype=np.float64)
                for a in (rho, ux, uy, uz)]
        _check(self._lib.lbm_init_equilibrium(self._h, *[abi.ptr(a) for a in arrs]), "lbm_init_equilibrium")

    def set_populations(self, f: np.ndarray):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(abi.Q, *self.shape)
        _check(self._lib.lbm_set_populations(self._h, abi.ptr(f)), "lbm_set_populations")

    def step(self, n: int = 1):
        _check(self._lib.lbm_step(self._h, int(n)), "lbm_step")

    def sync(self):
        _check(self._lib.lbm_sync(self._h), "lbm_sync")

    def probe(self, x: int, y: int, z: int):
        """(rho, ux, uy, uz) of one cell (caller coordinates x, y, z); synchronises."""
        out = np.empty(4)
        _check(self._lib.lbm_probe(self._h, int(x), int(y), int(z), abi.ptr(out)), "lbm_probe")
        return out

    def probe_record(self, x: int, y: int, z: int):
        """Record (rho, ux, uy, uz) of one cell on the device, without synchronising."""
        _check(self._lib.lbm_probe_record(self._h, int(x), int(y), int(z)), "lbm_probe_record")

    def probe_fetch(self):
        """The records made since the last fetch, shape (n, 4), oldest first."""
        n = ctypes.c_int64()
        _check(self._lib.lbm_probe_fetch(self._h, None, 0, ctypes.byref(n)), "lbm_probe_fetch")
        out = np.empty((n.value, 4))
        _check(self._lib.lbm_probe_fetch(self._h, abi.ptr(out), n.value, ctypes.byref(n)), "lbm_probe_fetch")
        return out

    def moment_add(self):
        """MomentTimeAvg::momentAdd of the current state (done by the next step's kernel)."""
        _check(self._lib.lbm_moment_add(self._h), "lbm_moment_add")

    def step_moment_add(self, n: int = 1):
        """n steps, adding rho and u of the state after each one to the time average."""
        _check(self._lib.lbm_step_moment_add(self._h, int(n)), "lbm_step_moment_add")

    def moment_avg(self):
        """MomentTimeAvg::momentAvg: (rho, ux, uy, uz) averages and the number of states added."""
        out = [np.empty(self.shape) for _ in range(4)]
        n = ctypes.c_int64()
        _check(self._lib.lbm_moment_avg(self._h, *[abi.ptr(a) for a in out], ctypes.byref(n)), "lbm_moment_avg")
        return out, n.value

    def moment_avg_reset(self):
        _check(self._lib.lbm_moment_avg_reset(self._h), "lbm_moment_avg_reset")

    def macroscopic(self):
        out = [np.empty(self.shape) for _ in range(4)]
        _check(self._lib.lbm_get_macroscopic(self._h, *[abi.ptr(a) for a in out]), "lbm_get_macroscopic")
        return out

    def populations(self):
        f = np.empty((abi.Q, *self.shape))
        _check(self._lib.lbm_get_populations(self._h, abi.ptr(f)), "lbm_get_populations")
        return f

    def omega(self):
        w = np.empty(self.shape)
        _check(self._lib.lbm_get_omega(self._h, abi.ptr(w)), "lbm_get_omega")
        return w

    def conservation(self):
        out = np.empty(4)
        _check(self._lib.lbm_conservation(self._h, abi.ptr(out)), "lbm_conservation")
        return out

    def step_diagnostics(self):
        """One step; returns (mass, momentum[3], wall force[3]) of that step."""
        out = np.empty(7)
        _check(self._lib.lbm_step_diagnostics(self._h, abi.ptr(out)), "lbm_step_diagnostics")
        return out[0], out[1:4].copy(), out[4:7].copy()

    def step_diagnostics_noalloc(self):
        """step_diagnostics through the entry point that allocates nothing."""
        out = np.empty(7)
        _check(self._lib.lbm_step_diagnostics_noalloc(self._h, abi.ptr(out)), "lbm_step_diagnostics_noalloc")
        return out[0], out[1:4].copy(), out[4:7].copy()

    def population(self, q: int):
        """Population plane q of the current state, shaped (nz, ny, nx)."""
        f = np.empty(self.shape)
        _check(self._lib.lbm_get_population(self._h, int(q), abi.ptr(f)), "lbm_get_population")
        return f

    def snapshot_begin(self):
        """Queue a snapshot of rho, u, omega behind the queued steps (returns at once)."""
        _check(self._lib.lbm_snapshot_begin(self._h), "lbm_snapshot_begin")

    def snapshot_wait(self):
        """(step, [rho, ux, uy, uz, omega]) of the pending snapshot."""
        out = [np.empty(self.shape) for _ in range(5)]
        st = ctypes.c_int64()
        _check(self._lib.lbm_snapshot_wait(self._h, *[abi.ptr(a) for a in out], ctypes.byref(st)),
               "lbm_snapshot_wait")
        return st.value, out

    def error_quadratic(self) -> float:
        """ErrorQuadratic's convergence norm against the previous call's
        velocity (reduced on the device); NaN sums raise LBM_E_NAN."""
        v = ctypes.c_double()
        _check(self._lib.lbm_error_quadratic(self._h, ctypes.byref(v)), "lbm_error_quadratic")
        return v.value

    def steps(self) -> int:
        v = ctypes.c_int64()
        _check(self._lib.lbm_step_count(self._h, ctypes.byref(v)), "lbm_step_count")
        return v.value

    def last_kernel_ms(self) -> float:
        v = ctypes.c_double()
        _check(self._lib.lbm_last_kernel_ms(self._h, ctypes.byref(v)), "lbm_last_kernel_ms")
        return v.value

    def device_bytes(self) -> int:
        v = ctypes.c_size_t()
        _check(self._lib.lbm_device_bytes(self._h, ctypes.byref(v)), "lbm_device_bytes")
        return v.value

    # -- learned closure and filtering -------------------------------------
    def set_closure(self, params: np.ndarray, norm: np.ndarray):
        """Install closure weights (NPARAM floats) and normalisation
        {mu, sd, lo, hi} (4 x 12 floats); lattice created with CLOSURE_MLP."""
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        nm = np.ascontiguousarray(norm, dtype=np.float32).ravel()
        if p.size != abi.MLP_PARAMS or nm.size != 4 * abi.MLP_FEATURES:
            raise ValueError("set_closure: wrong parameter or normalisation size")
        _check(self._lib.lbm_set_closure(self._h, abi.ptr(p, ctypes.c_float), abi.ptr(nm, ctypes.c_float)),
               "lbm_set_closure")

    def set_closure_refresh(self, every: int):
        """Evaluate the closure network every `every` steps, holding C(x) between."""
        _check(self._lib.lbm_set_closure_refresh(self._h, int(every)), "lbm_set_closure_refresh")

    def closure_features(self):
        """(normalised features [12, nz, ny, nx], envelope flag, coefficient)."""
        f = np.empty((abi.MLP_FEATURES, *self.shape), np.float32)
        e = np.empty(self.shape, np.int32)
        c = np.empty(self.shape)
        _check(self._lib.lbm_get_closure_features(self._h, abi.ptr(f, ctypes.c_float), abi.ptr(e, ctypes.c_int32),
                                                  abi.ptr(c)), "lbm_get_closure_features")
        return f, e.astype(bool), c

    def filter(self, factor: int):
        """Box filter over factor^3 blocks: 10 arrays (rho, ux, uy, uz, uxux,
        uyuy, uzuz, uxuy, uxuz, uyuz) shaped (nz/f, ny/f, nx/f)."""
        nz, ny, nx = self.shape
        cs = (nz // factor, ny // factor, nx // factor)
        out = np.empty((10, *cs))
        _check(self._lib.lbm_filter(self._h, int(factor), abi.ptr(out)), "lbm_filter")
        return out
