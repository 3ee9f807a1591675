"""Seeded initial fields for the benchmark configurations (host side, numpy).

* decaying_turbulence: periodic solenoidal velocity from random Fourier modes
  with E(k) ~ k^4 exp(-2 (k/k0)^2), uniform random phases, Hermitian symmetry
  (real inverse FFT) and rms velocity u_rms = sqrt(<u.u>/3).
* taylor_green: the Taylor-Green vortex with its pressure-consistent density.
* energy_spectrum: shell-summed E(k) of a periodic velocity field (Parseval:
  sum_k E(k) = <u.u>/2).

Arrays are shaped (nz, ny, nx) with x fastest, like the lattice.
"""
from __future__ import annotations

import numpy as np


def _wavenumbers(n):
    return np.fft.fftfreq(n, d=1.0 / n)


def decaying_turbulence(n: int, k0: float = 4.0, u_rms: float = 0.04, seed: int = 20240601,
                        dtype=np.float64):
    """Velocity (ux, uy, uz) on an n^3 periodic grid, E(k) ~ k^4 exp(-2 (k/k0)^2)."""
    rng = np.random.default_rng(seed)
    kz = _wavenumbers(n)[:, None, None]
    ky = _wavenumbers(n)[None, :, None]
    kx = np.fft.rfftfreq(n, d=1.0 / n)[None, None, :]
    k2 = kz * kz + ky * ky + kx * kx
    k = np.sqrt(k2)
    with np.errstate(divide="ignore", invalid="ignore"):
        spec = k ** 4 * np.exp(-2.0 * (k / k0) ** 2)
        # shell density: |u_hat(k)|^2 ~ E(k) / (4 pi k^2)
        amp = np.where(k > 0, np.sqrt(spec / (4.0 * np.pi * k2)), 0.0)
    shape = k.shape
    comps = []
    for _ in range(3):
        phase = rng.uniform(0.0, 2.0 * np.pi, shape)
        mag = rng.standard_normal(shape)
        comps.append(amp * mag * np.exp(1j * phase))
    ux, uy, uz = comps
    # solenoidal projection: u_hat <- u_hat - k (k . u_hat) / k^2
    with np.errstate(divide="ignore", invalid="ignore"):
        kdotu = (kx * ux + ky * uy + kz * uz) / np.where(k2 > 0, k2, 1.0)
    ux = ux - kx * kdotu
    uy = uy - ky * kdotu
    uz = uz - kz * kdotu
    out = [np.fft.irfftn(c, s=(n, n, n), axes=(0, 1, 2)) for c in (ux, uy, uz)]
    rms = np.sqrt(sum(np.mean(c * c) for c in out) / 3.0)
    return tuple((c * (u_rms / rms)).astype(dtype) for c in out)


def taylor_green(n: int, u0: float = 0.05):
    """Taylor-Green vortex on an n^3 periodic grid, x = 2 pi (i + 1/2) / n.

    u = u0 sin x cos y cos z, v = -u0 cos x sin y cos z, w = 0,
    rho = 1 + 3 (u0^2 / 16) (cos 2x + cos 2y)(cos 2z + 2)  (p = rho / 3)."""
    c = 2.0 * np.pi * (np.arange(n) + 0.5) / n
    z = c[:, None, None]
    y = c[None, :, None]
    x = c[None, None, :]
    ux = u0 * np.sin(x) * np.cos(y) * np.cos(z)
    uy = -u0 * np.cos(x) * np.sin(y) * np.cos(z)
    uz = np.zeros_like(ux * 0 + z * 0)
    rho = 1.0 + 3.0 * (u0 * u0 / 16.0) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    shape = (n, n, n)
    return (np.broadcast_to(rho, shape).copy(), np.broadcast_to(ux, shape).copy(),
            np.broadcast_to(uy, shape).copy(), np.broadcast_to(uz, shape).copy())


def energy_spectrum(ux, uy, uz):
    """Shell-summed kinetic-energy spectrum of a periodic field.

    Returns (k, E) with E[k] the energy in shells |k| in [k - 1/2, k + 1/2);
    sum(E) equals mean(u.u)/2 (Parseval)."""
    n = ux.shape[0]
    N = ux.size
    e = np.zeros(ux.shape[:2] + (ux.shape[2] // 2 + 1,))
    for c in (ux, uy, uz):
        h = np.fft.rfftn(c) / N
        e += np.abs(h) ** 2
    # rfft stores half the spectrum: double every column except k_x = 0 (and Nyquist for even n)
    w = np.full(e.shape[2], 2.0)
    w[0] = 1.0
    if ux.shape[2] % 2 == 0:
        w[-1] = 1.0
    e *= w[None, None, :] * 0.5
    kz = _wavenumbers(ux.shape[0])[:, None, None]
    ky = _wavenumbers(ux.shape[1])[None, :, None]
    kx = np.fft.rfftfreq(ux.shape[2], d=1.0 / ux.shape[2])[None, None, :]
    kk = np.rint(np.sqrt(kz * kz + ky * ky + kx * kx)).astype(np.int64)
    E = np.bincount(kk.ravel(), weights=e.ravel())
    return np.arange(E.size), E


def divergence_spectral(ux, uy, uz):
    """RMS of the spectral divergence of a periodic field."""
    kz = _wavenumbers(ux.shape[0])[:, None, None]
    ky = _wavenumbers(ux.shape[1])[None, :, None]
    kx = _wavenumbers(ux.shape[2])[None, None, :]
    d = 1j * (kx * np.fft.fftn(ux) + ky * np.fft.fftn(uy) + kz * np.fft.fftn(uz))
    return float(np.sqrt(np.mean(np.abs(np.fft.ifftn(d)) ** 2)))
