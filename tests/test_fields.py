"""Seeded initial fields and the energy spectrum (host, numpy).

- Taylor-Green: the closed form, divergence-free, with the density that
  balances its pressure.
- Decaying turbulence: rms velocity as requested, divergence-free to FFT
  rounding, seeded (same seed -> same field), and its measured shell spectrum
  follows the prescribed k^4 exp(-2 (k/k0)^2) shape.
- energy_spectrum: Parseval (sum of E(k) = <u.u>/2) and a single Fourier mode
  lands in its own shell with the mode's energy.
- dissipation: -dK/dt of an exponential decay against its closed form.
"""
import numpy as np

from lb_amd import closure
from lb_amd.fields import decaying_turbulence, divergence_spectral, energy_spectrum, taylor_green


def test_taylor_green_closed_form():
    n, u0 = 32, 0.05
    rho, ux, uy, uz = taylor_green(n, u0)
    c = 2 * np.pi * (np.arange(n) + 0.5) / n
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    assert np.allclose(ux, u0 * np.sin(x) * np.cos(y) * np.cos(z), atol=1e-15)
    assert np.allclose(uy, -u0 * np.cos(x) * np.sin(y) * np.cos(z), atol=1e-15)
    assert np.all(uz == 0)
    assert np.allclose(rho, 1 + 3 * u0 * u0 / 16 * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2), atol=1e-15)
    assert divergence_spectral(ux, uy, uz) < 1e-15
    # a single mode per component: all energy in the shell |k| = sqrt(3)
    k, E = energy_spectrum(ux, uy, uz)
    assert np.isclose(E.sum(), 0.5 * np.mean(ux * ux + uy * uy + uz * uz), rtol=1e-12)
    assert np.isclose(E[2], E.sum(), rtol=1e-12)


def test_decaying_turbulence_properties():
    n, k0, urms = 48, 4.0, 0.04
    ux, uy, uz = decaying_turbulence(n, k0=k0, u_rms=urms, seed=7)
    assert np.isclose(np.sqrt(np.mean(ux * ux + uy * uy + uz * uz) / 3), urms, rtol=1e-12)
    assert divergence_spectral(ux, uy, uz) < 1e-12 * urms * n
    again = decaying_turbulence(n, k0=k0, u_rms=urms, seed=7)
    other = decaying_turbulence(n, k0=k0, u_rms=urms, seed=8)
    assert all(np.array_equal(a, b) for a, b in zip((ux, uy, uz), again))
    assert not np.array_equal(ux, other[0])
    # measured shell spectrum against the prescribed one, summed over the
    # same lattice shells (the spectrum falls steeply across a shell, so its
    # value at the shell's integer k is not the shell's energy): E(k) / (4 pi k^2)
    # per mode.  The random amplitudes scatter each shell by about
    # sqrt(2 / modes), 5-15 % for shells 2..12.
    k, E = energy_spectrum(ux, uy, uz)
    f = np.fft.fftfreq(n, d=1.0 / n)
    kz, ky, kx = np.meshgrid(f, f, f, indexing="ij")
    km = np.sqrt(kx * kx + ky * ky + kz * kz)
    with np.errstate(divide="ignore", invalid="ignore"):
        per_mode = np.where(km > 0, km ** 4 * np.exp(-2 * (km / k0) ** 2) / (4 * np.pi * km * km), 0.0)
    model = np.bincount(np.rint(km).astype(np.int64).ravel(), weights=per_mode.ravel())
    ratio = E[2:13] / model[2:13]
    ratio /= ratio.mean()
    assert np.std(ratio) < 0.15 and np.all(np.abs(ratio - 1) < 0.35), ratio
    assert int(np.argmax(E)) in (3, 4, 5)   # k^4 exp(-2 (k/k0)^2) peaks at k0 = 4


def test_spectrum_parseval_and_single_mode():
    n = 24
    rng = np.random.default_rng(3)
    u = [rng.standard_normal((n, n, n)) for _ in range(3)]
    k, E = energy_spectrum(*u)
    assert np.isclose(E.sum(), 0.5 * np.mean(u[0] ** 2 + u[1] ** 2 + u[2] ** 2), rtol=1e-12)
    # one mode of wavenumber (3, 4, 0): |k| = 5
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    ux = 0.1 * np.cos(2 * np.pi * (3 * x + 4 * y) / n)
    zero = np.zeros_like(ux)
    k, E = energy_spectrum(ux, zero, zero)
    assert np.isclose(E[5], 0.5 * np.mean(ux * ux), rtol=1e-12)
    assert np.isclose(E.sum(), E[5], rtol=1e-12)


def test_dissipation_of_exponential_decay():
    t = np.linspace(0.0, 100.0, 201)
    lam, K0 = 0.013, 2.5e-3
    K = K0 * np.exp(-lam * t)
    eps = closure.dissipation(t, K)
    exact = lam * K
    # central differences: relative error (lam dt)^2 / 6 in the interior
    assert np.allclose(eps[1:-1], exact[1:-1], rtol=(lam * 0.5) ** 2)
