"""Size-independent properties at the full workload sizes (the oracle is too
slow there): the decaying-turbulence lattice at 256^3, the Taylor-Green
lattice at 384^3 and the obstacle workload's 512 x 256 x 256 shape.

- Translation invariance: every cell runs the same arithmetic, so on an
  index-wrap periodic lattice a shifted initial field gives the shifted
  result bit for bit.  This exercises the tile mapping, the halo layout, the
  wrap deltas and the 32-bit offsets over the whole lattice, on even and odd
  steps.
- Conservation: with the true opposite table and exact weights (compat = 0)
  FP64 storage keeps mass and momentum to rounding.
"""
import numpy as np
import pytest

from lb_amd import Lattice, abi, make_desc
from lb_amd.fields import decaying_turbulence, taylor_green

pytestmark = pytest.mark.gpu

SHIFT = (3, 2, 1)   # (z, y, x)


def _lattice(n, storage, collision=abi.COLLISION_CUM2015, **kw):
    flags = np.ones((n, n, n), np.int32)
    return Lattice(make_desc(flags, [], collision=collision, closure=abi.CLOSURE_SMAGORINSKY,
                             c_les=0.17, storage=storage, arith=abi.ARITH_FAST, periodic=(1, 1, 1), **kw))


def _shifted_runs(n, storage, fields, steps, **kw):
    out = []
    for shift in ((0, 0, 0), SHIFT):
        lat = _lattice(n, storage, **kw)
        init = [None if a is None else np.roll(a, shift, axis=(0, 1, 2)) for a in fields]
        lat.init_equilibrium(*init)
        lat.step(steps)
        out.append(lat.macroscopic())
        lat.close()
    return out


@pytest.mark.parametrize("storage", [abi.STORAGE_F32, abi.STORAGE_F16DEV], ids=["f32", "f16dev"])
def test_translation_invariance_256(gpu, storage):
    n = 256
    ux, uy, uz = decaying_turbulence(n, k0=4.0, u_rms=0.04, seed=20240601)
    a, b = _shifted_runs(n, storage, (None, ux, uy, uz), 7, compat=abi.COMPAT_REFERENCE, tau=0.5015)
    for k in range(4):
        assert np.array_equal(np.roll(a[k], SHIFT, axis=(0, 1, 2)), b[k]), k
    assert np.all(np.isfinite(a[1]))


def test_translation_invariance_384_taylor_green(gpu):
    n = 384
    rho, ux, uy, uz = taylor_green(n, 0.05)
    nu = 0.05 * (n / (2 * np.pi)) / 1600.0
    a, b = _shifted_runs(n, abi.STORAGE_BF16DEV, (rho, ux, uy, uz), 5, compat=0, tau=3 * nu + 0.5)
    for k in range(4):
        assert np.array_equal(np.roll(a[k], SHIFT, axis=(0, 1, 2)), b[k]), k


def test_translation_invariance_haloed_512x256x256(gpu):
    """The obstacle workload's lattice shape, 512 x 256 x 256 with a solid halo
    along x (whole-cache-line x halo, boundary cells at both x ends): shifts
    along the wrapped y and z axes give the shifted result bit for bit."""
    nx, ny, nz = 512, 256, 256
    flags = np.ones((nz, ny, nx), np.int32)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij", sparse=True)
    ux = 0.03 * np.sin(2 * np.pi * y / ny) * np.cos(2 * np.pi * z / nz) * np.sin(np.pi * (x + 0.5) / nx)
    uy = 0.02 * np.cos(2 * np.pi * x / nx) * np.ones((nz, ny, nx))
    out = []
    for shift in ((0, 0), (5, 3)):
        lat = Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY,
                                c_les=0.17, storage=abi.STORAGE_F32, arith=abi.ARITH_FAST, compat=0,
                                periodic=(0, 1, 1), tau=0.52))
        lat.init_equilibrium(None, np.roll(ux, shift, axis=(0, 1)), np.roll(uy, shift, axis=(0, 1)), None)
        lat.step(6)
        out.append(lat.macroscopic())
        lat.close()
    for k in range(4):
        assert np.array_equal(np.roll(out[0][k], (5, 3), axis=(0, 1)), out[1][k]), k
    assert np.all(np.isfinite(out[0][1]))


def test_conservation_fp64_256(gpu):
    n = 256
    ux, uy, uz = decaying_turbulence(n, k0=4.0, u_rms=0.04, seed=20240601)
    lat = _lattice(n, abi.STORAGE_F64, compat=0, tau=0.5015)
    lat.init_equilibrium(None, ux, uy, uz)
    lat.step(1)
    m0 = lat.conservation()
    lat.step(10)
    m1 = lat.conservation()
    print(f"[conservation 256^3 f64] mass {m0[0]:.17g} -> {m1[0]:.17g}, "
          f"momentum drift {np.abs(m1[1:] - m0[1:]).max():.3e}", flush=True)
    assert abs(m1[0] - m0[0]) <= 1e-12 * m0[0]
    assert np.all(np.abs(m1[1:] - m0[1:]) <= 1e-12 * m0[0])


@pytest.mark.parametrize("storage", [abi.STORAGE_F32, abi.STORAGE_F16DEV], ids=["f32", "f16dev"])
def test_translation_invariance_256_cumulant_2017(gpu, storage):
    """The 2017 operator's limiters and rate functions act per cell: shifted
    in, shifted out, bit for bit."""
    n = 256
    ux, uy, uz = decaying_turbulence(n, k0=4.0, u_rms=0.04, seed=20240602)
    a, b = _shifted_runs(n, storage, (None, ux, uy, uz), 6, compat=abi.COMPAT_REFERENCE, tau=0.5015,
                         collision=abi.COLLISION_CUM2017)
    for k in range(4):
        assert np.array_equal(np.roll(a[k], SHIFT, axis=(0, 1, 2)), b[k]), k


def test_conservation_fp64_256_cumulant_2017(gpu):
    n = 256
    ux, uy, uz = decaying_turbulence(n, k0=4.0, u_rms=0.04, seed=20240603)
    lat = Lattice(make_desc(np.ones((n, n, n), np.int32), [], collision=abi.COLLISION_CUM2017,
                            closure=abi.CLOSURE_SMAGORINSKY, c_les=0.17, storage=abi.STORAGE_F64,
                            arith=abi.ARITH_FAST, compat=0, periodic=(1, 1, 1), tau=0.5015))
    lat.init_equilibrium(None, ux, uy, uz)
    lat.step(1)
    m0 = lat.conservation()
    lat.step(10)
    m1 = lat.conservation()
    print(f"[conservation 256^3 f64 cum2017] mass drift {abs(m1[0] - m0[0]):.3e}, "
          f"momentum drift {np.abs(m1[1:] - m0[1:]).max():.3e}", flush=True)
    assert abs(m1[0] - m0[0]) <= 1e-12 * m0[0]
    assert np.all(np.abs(m1[1:] - m0[1:]) <= 1e-12 * m0[0])
