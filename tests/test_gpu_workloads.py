"""The obstacle and Taylor-Green workloads with their real geometry and
storage formats.

- Obstacle scene (synthetic OFF: domain box, seeded sphere and ellipsoids),
  voxelized by the host and GPU voxelizers, meshed by the host boundary
  mesher: equilibrium velocity inlet at x = 0 (InletVelocityEquilibrium.h:
  51-103), interpolated density outlet at x = nx - 1
  (OutletDensityInterpolated.h:170-252), walls on every obstacle, index wrap
  along y and z, cumulant + Smagorinsky.
  * a 64 x 32 x 32 scene against the oracle, FP64 literal, bit for bit;
  * the 512 x 256 x 256 scene: the flag grid, the boundary list and the
    initial field shifted along the wrapped axes give the shifted result bit
    for bit (every boundary rule runs per cell or per list entry), with a
    finite state and drag.
- Taylor-Green 384^3, Re 1600: FP16-deviation storage, with nearest and
  stochastic rounding, against FP64 storage after 200 steps.
"""
import tempfile

import numpy as np
import pytest

from helpers import OracleLattice, rel_l2
from lb_amd import Lattice, abi, make_desc
from lb_amd.fields import taylor_green
from lb_amd.geometry import obstacle_scene
from lb_amd.studies import c4_geometry

pytestmark = pytest.mark.gpu

U_IN = 0.04


def _scene(small):
    with tempfile.TemporaryDirectory() as td:
        if small:
            scene = obstacle_scene(domain=(64.0, 32.0, 32.0), sphere_radius=4.0, n_theta=10, n_phi=14,
                                   ellipsoid_radii=((2, 2, 2), (5, 4, 4)), margin=2.0)
            return c4_geometry(td, scene=scene, u_in=U_IN)
        return c4_geometry(td, u_in=U_IN)


@pytest.mark.parametrize("compat", [abi.COMPAT_REFERENCE, 0], ids=["reference", "corrected"])
def test_obstacle_scene_64x32x32_equals_oracle(gpu, compat):
    flags, bc, info = _scene(True)
    kinds = {b.kind for b in bc}
    assert {abi.BC_WALL, abi.BC_INLET_EQUILIBRIUM, abi.BC_OUTLET_INTERPOLATED} <= kinds
    assert len(info["obstacles"]) == 4 and info["analytic_mismatch_far"] == 0
    nu = U_IN * 8.0 / 100.0
    kw = dict(collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY, c_les=0.17, storage=abi.STORAGE_F64,
              arith=abi.ARITH_LITERAL, compat=compat, periodic=(0, 1, 1), tau=3 * nu + 0.5, nu=nu)
    g = Lattice(make_desc(flags, bc, **kw))
    o = OracleLattice(make_desc(flags, bc, **kw))
    ux0 = np.where(flags == 0, 0.0, U_IN)
    for L in (g, o):
        L.init_equilibrium(None, ux0, None, None)
    g.step(60)
    o.step(60)
    fluid = flags != 0
    fg, fo = g.populations(), o.populations()
    mg, mo_ = g.macroscopic(), o.macroscopic()
    ef = rel_l2(fg[:, fluid], fo[:, fluid])
    eu = rel_l2(np.stack([m[fluid] for m in mg[1:]]), np.stack([m[fluid] for m in mo_[1:]]))
    print(f"[obstacle 64x32x32 compat={compat}] relL2 f={ef:.3e} u={eu:.3e} "
          f"walls={info['wall_nodes']}", flush=True)
    assert ef <= 1e-12 and eu <= 1e-12
    assert np.abs(mg[1][fluid]).max() > 0.5 * U_IN


def test_obstacle_scene_512x256x256_translation(gpu):
    flags, bc, info = _scene(False)
    nz, ny, nx = flags.shape
    assert (nx, ny, nz) == (512, 256, 256) and info["triangles"] > 19000
    nu = U_IN * 48.0 / 1000.0
    shift = (7, 5)   # (z, y)
    out, drag = [], []
    for sz, sy in ((0, 0), shift):
        fl = np.roll(flags, (sz, sy), axis=(0, 1))
        bcs = []
        for b in bc:
            c = type(b)()
            ctype = type(b)
            for name, _ in ctype._fields_:
                setattr(c, name, getattr(b, name))
            c.y = (b.y + sy) % ny
            c.z = (b.z + sz) % nz
            bcs.append(c)
        lat = Lattice(make_desc(fl, bcs, collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY,
                                c_les=0.17, storage=abi.STORAGE_F32, arith=abi.ARITH_FAST, compat=abi.COMPAT_REFERENCE,
                                periodic=(0, 1, 1), tau=3 * nu + 0.5, nu=nu))
        lat.init_equilibrium(None, np.where(fl == 0, 0.0, U_IN), None, None)
        lat.step(5)
        _, _, force = lat.step_diagnostics()
        drag.append(force)
        out.append(lat.macroscopic())
        lat.close()
    for k in range(4):
        assert np.array_equal(np.roll(out[0][k], shift, axis=(0, 1)), out[1][k]), k
    fluid = flags != 0
    assert all(np.all(np.isfinite(m[fluid])) for m in out[0])
    print(f"[obstacle 512x256x256] drag {drag[0][0]:.6e} / shifted {drag[1][0]:.6e}, "
          f"walls {info['wall_nodes']}", flush=True)
    # the obstacles take momentum out of the flow along +x
    assert drag[0][0] > 0 and abs(drag[1][0] - drag[0][0]) <= 1e-9 * abs(drag[0][0])


def test_taylor_green_384_fp16_storage_against_fp64(gpu):
    n = 384
    rho, ux, uy, uz = taylor_green(n, 0.05)
    nu = 0.05 * (n / (2 * np.pi)) / 1600.0
    kw = dict(collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_NONE, arith=abi.ARITH_FAST, compat=0,
              periodic=(1, 1, 1), tau=3 * nu + 0.5)
    flags = np.ones((n, n, n), np.int32)
    steps = 200
    ref = Lattice(make_desc(flags, [], storage=abi.STORAGE_F64, **kw))
    ref.init_equilibrium(rho, ux, uy, uz)
    ref.step(steps)
    mr = ref.macroscopic()
    ref.close()
    ur = np.stack(mr[1:])
    bounds = {abi.STORAGE_F16DEV: 4e-3, abi.STORAGE_F16DEV_SR: 5e-4}
    for st, bound in bounds.items():
        lat = Lattice(make_desc(flags, [], storage=st, **kw))
        lat.init_equilibrium(rho, ux, uy, uz)
        lat.step(steps)
        m = lat.macroscopic()
        lat.close()
        eu = rel_l2(np.stack(m[1:]), ur)
        er = rel_l2(m[0] - 1.0, mr[0] - 1.0)
        print(f"[taylor-green 384^3 {abi.STORAGE_NAMES[st]} vs f64, {steps} steps] relL2 u {eu:.3e}, "
              f"rho-1 {er:.3e}", flush=True)
        assert eu <= bound
        del m
