"""Viscous decay on the GPU against its closed form.

A small-amplitude Taylor-Green vortex on a periodic 64^3 lattice stays in the
linear regime: each velocity component is one Fourier mode with |k|^2 =
3 (2 pi / n)^2, so the kinetic energy decays as K0 exp(-2 nu |k|^2 t) with the
lattice viscosity nu = (tau - 1/2) / 3.  This pins the collision's viscosity
(both operators) and the dissipation estimate eps = -dK/dt used by the
closure experiment, against eps = 2 nu |k|^2 K.
"""
import numpy as np
import pytest

from lb_amd import Lattice, abi, closure, make_desc
from lb_amd.fields import taylor_green

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("coll", [abi.COLLISION_CUM2015, abi.COLLISION_SRT], ids=["cum2015", "srt"])
def test_taylor_green_linear_decay_rate(gpu, coll):
    n, u0, tau = 64, 1e-3, 0.6
    nu = (tau - 0.5) / 3.0
    k2 = 3.0 * (2.0 * np.pi / n) ** 2
    rho, ux, uy, uz = taylor_green(n, u0)
    flags = np.ones((n, n, n), np.int32)
    lat = Lattice(make_desc(flags, [], collision=coll, storage=abi.STORAGE_F64, arith=abi.ARITH_FAST, compat=0,
                            periodic=(1, 1, 1), tau=tau))
    lat.init_equilibrium(rho, ux, uy, uz)
    t, K = [], []
    every = 100
    for s in range(0, 1201, every):
        if s:
            lat.step(every)
        m = lat.macroscopic()
        t.append(s)
        K.append(0.5 * float(np.mean(m[1] ** 2 + m[2] ** 2 + m[3] ** 2)))
    t, K = np.array(t, float), np.array(K)
    # fitted decay rate of K after the first sample (initial equilibrium start)
    rate = -np.polyfit(t[1:], np.log(K[1:]), 1)[0]
    exact = 2.0 * nu * k2
    eps = closure.dissipation(t, K)[2:-1]
    eps_exact = exact * K[2:-1]
    print(f"[linear decay coll={coll}] rate {rate:.6e} exact {exact:.6e} ({rate / exact - 1:+.2e}); "
          f"eps max rel err {np.abs(eps / eps_exact - 1).max():.2e}", flush=True)
    assert abs(rate / exact - 1) < 5e-3
    assert np.allclose(eps, eps_exact, rtol=2e-2)
