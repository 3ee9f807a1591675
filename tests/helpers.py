"""Test-side helpers: ctypes bindings of the CPU oracle and geometry builders.

Only tests, __graft_entry__.smoke() and bench.py's CPU-baseline leg use the
oracle; it is the checker, never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from lb_amd import abi  # noqa: E402
from oracle import mesher_oracle  # noqa: E402

ORACLE_LIB = os.path.join(ROOT, "oracle", "liblbm_oracle.so")
VOXEL_LIB = os.path.join(ROOT, "oracle", "libvoxel_oracle.so")
HOSTCOLL_LIB = os.path.join(ROOT, "tests", "native", "libcollide_host.so")

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def _load(path, what):
    if not os.path.exists(path):
        raise RuntimeError(f"{what} {path} is not built; run __graft_entry__.build()")
    return ctypes.CDLL(path)


_orc = None


def oracle_lib():
    global _orc
    if _orc is None:
        L = _load(ORACLE_LIB, "oracle")
        sig = {
            "orc_create": (ctypes.c_int, [ctypes.POINTER(abi.Desc), ctypes.POINTER(_vp)]),
            "orc_init_equilibrium": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp]),
            "orc_set_populations": (ctypes.c_int, [_vp, _dp]),
            "orc_step": (ctypes.c_int, [_vp, ctypes.c_int]),
            "orc_get_macroscopic": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp]),
            "orc_moment_add": (ctypes.c_int, [_vp]),
            "orc_moment_avg": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_longlong)]),
            "orc_get_populations": (ctypes.c_int, [_vp, _dp]),
            "orc_get_post_populations": (ctypes.c_int, [_vp, _dp]),
            "orc_get_omega": (ctypes.c_int, [_vp, _dp]),
            "orc_error_quadratic": (ctypes.c_int, [_vp, _dp]),
            "orc_destroy": (None, [_vp]),
            "orc_set_threads": (None, [ctypes.c_int]),
            "orc_get_threads": (ctypes.c_int, []),
            "orc_cell_collide": (None, [ctypes.c_int, ctypes.c_uint32, ctypes.c_int, _dp, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, _dp]),
            "orc_cell_omega_les": (ctypes.c_double, [ctypes.c_uint32, ctypes.c_int, _dp, ctypes.c_double,
                                                     ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                     ctypes.c_double, ctypes.c_double]),
            "orc_cell_feq": (None, [ctypes.c_uint32, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, _dp]),
            "orc_tables": (None, [ctypes.c_uint32, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _dp]),
            "orc_cell_central_moments": (None, [_dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]),
            "orc_cell_from_central_moments": (None, [_dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]),
        }
        for k, (r, a) in sig.items():
            fn = getattr(L, k)
            fn.restype, fn.argtypes = r, a
        _orc = L
    return _orc


_hc = None


def host_collide_lib():
    global _hc
    if _hc is None:
        L = _load(HOSTCOLL_LIB, "host collide harness")
        L.host_collide.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _dp, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.host_collide.restype = None
        L.host_moments.argtypes = [ctypes.c_int, _dp, _dp]
        L.host_moments.restype = None
        L.host_omega_les.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _dp, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double]
        L.host_omega_les.restype = ctypes.c_double
        L.host_strain_features.argtypes = [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double, _dp]
        L.host_strain_features.restype = None
        L.host_cell_features.argtypes = [_dp, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        L.host_cell_features.restype = None
        _hc = L
    return _hc


def tables(compat=abi.COMPAT_REFERENCE):
    c = np.zeros(81, np.int32)
    cr = np.zeros(27, np.int32)
    w = np.zeros(27)
    oracle_lib().orc_tables(compat, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            cr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), abi.ptr(w))
    return c.reshape(27, 3), cr, w


class OracleLattice:
    """Same interface as lb_amd.Lattice, backed by the CPU oracle."""

    def __init__(self, desc):
        self._lib = oracle_lib()
        self.desc = desc
        self.shape = (desc.nz, desc.ny, desc.nx)
        h = ctypes.c_void_p()
        st = self._lib.orc_create(ctypes.byref(desc), ctypes.byref(h))
        if st != 0:
            raise RuntimeError(f"orc_create failed: {abi.STATUS_NAMES.get(st, st)}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_equilibrium(self, rho=None, ux=None, uy=None, uz=None):
        arrs = [None if a is None else np.ascontiguousarray(np.broadcast_to(a, self.shape), dtype=np.float64)
                for a in (rho, ux, uy, uz)]
        assert self._lib.orc_init_equilibrium(self._h, *[abi.ptr(a) for a in arrs]) == 0

    def set_populations(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert self._lib.orc_set_populations(self._h, abi.ptr(f)) == 0

    def step(self, n=1):
        assert self._lib.orc_step(self._h, int(n)) == 0

    def macroscopic(self):
        out = [np.empty(self.shape) for _ in range(4)]
        assert self._lib.orc_get_macroscopic(self._h, *[abi.ptr(a) for a in out]) == 0
        return out

    def populations(self):
        f = np.empty((27, *self.shape))
        assert self._lib.orc_get_populations(self._h, abi.ptr(f)) == 0
        return f

    def moment_add(self):
        assert self._lib.orc_moment_add(self._h) == 0

    def step_moment_add(self, n=1):
        for _ in range(int(n)):
            self.step(1)
            self.moment_add()

    def moment_avg(self):
        out = [np.empty(self.shape) for _ in range(4)]
        n = ctypes.c_longlong()
        assert self._lib.orc_moment_avg(self._h, *[abi.ptr(a) for a in out], ctypes.byref(n)) == 0
        return out, n.value

    def post_populations(self):
        f = np.empty((27, *self.shape))
        assert self._lib.orc_get_post_populations(self._h, abi.ptr(f)) == 0
        return f

    def omega(self):
        w = np.empty(self.shape)
        assert self._lib.orc_get_omega(self._h, abi.ptr(w)) == 0
        return w

    def error_quadratic(self):
        e = ctypes.c_double()
        assert self._lib.orc_error_quadratic(self._h, ctypes.byref(e)) == 0
        return e.value


def bc_nodes(bc):
    """List of dicts (oracle.mesher_oracle) -> list of abi.BcNode."""
    out = []
    for d in bc:
        n = abi.BcNode()
        n.x, n.y, n.z, n.kind = d["x"], d["y"], d["z"], d["kind"]
        nrm = d.get("normal", (0.0, 0.0, 0.0))
        vel = d.get("velocity", (0.0, 0.0, 0.0))
        n.nx, n.ny, n.nz = nrm
        n.ux, n.uy, n.uz = vel
        n.rho = d.get("density", 0.0)
        n.per_idx = d.get("per_idx", 0)
        n.regular = d.get("regular", 1)
        out.append(n)
    return out


def cavity_setup(n, lid=(0.05, 0.0, 0.0)):
    """Flags, BcNode list and initial u (inlet velocity on lid cells) of the
    lid-driven cavity with n^3 fluid cells."""
    flags, bc = mesher_oracle.cavity(n, lid)
    ux = np.zeros(flags.shape)
    for d in bc:
        if d["kind"] == mesher_oracle.BC_INLET_MOVING_WALL:
            ux[d["z"], d["y"], d["x"]] = d["velocity"][0]
    return flags, bc_nodes(bc), ux


def rel_l2(a, b, mask=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if mask is not None:
        a, b = a[mask], b[mask]
    den = np.sqrt(np.sum(b * b))
    num = np.sqrt(np.sum((a - b) ** 2))
    return num / den if den > 0 else num


def random_velocity_field(shape, amp, seed, kmax=2):
    """Smooth periodic random velocity: a few low Fourier modes per component."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    out = []
    for _ in range(3):
        f = np.zeros(shape)
        for kx in range(0, kmax + 1):
            for ky in range(0, kmax + 1):
                for kz in range(0, kmax + 1):
                    if kx == ky == kz == 0:
                        continue
                    ph = rng.uniform(0, 2 * np.pi)
                    a = rng.standard_normal()
                    f += a * np.cos(2 * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz) + ph)
        f *= amp / max(np.abs(f).max(), 1e-30)
        out.append(f)
    return out
