"""Host-side setup library (no GPU): voxelizer, flag-grid files, boundary
mesher, unit conversion.  Checked against the reference's committed
voxelization, the oracle mesher / brute-force voxelizer, and analytic shapes."""
import gzip
import math
import os

import numpy as np
import pytest

from helpers import ROOT
from lb_amd import abi, host
from lb_amd.geometry import edge_manifold, mesh_volume, obstacle_scene, uv_ellipsoid, write_off
from oracle import mesher_oracle as mo
from test_voxelizer import GOLD_OFF, GOLD_TXT, OUTSIDE, golden_grid, oracle_voxelize


def test_voxelizer_matches_reference_grid_bit_exact():
    dims, gold = golden_grid()
    g, mn, mx = host.voxelize_off(GOLD_OFF, 0.25, 1.0, OUTSIDE)
    assert list(g.shape[::-1]) == dims
    assert mn == (0.0, 0.0, 0.0) and mx == (1500.0, 200.0, 200.0)
    assert np.array_equal(g, gold)


def test_mesh_file_is_byte_identical_to_reference(tmp_path):
    g, mn, mx = host.voxelize_off(GOLD_OFF, 0.25, 1.0, OUTSIDE)
    p = str(tmp_path / "BackwardStepLaminar.txt")
    host.mesh_write(p, g, 0.25, 1.0, mn, mx)
    assert open(p, "rb").read() == gzip.open(GOLD_TXT, "rb").read()
    g2, res, add, mn2, mx2 = host.mesh_read(p)
    assert np.array_equal(g2, g) and res == 0.25 and add == 1.0 and mn2 == mn and mx2 == mx


def test_mesh_read_rejects_truncated_and_missing(tmp_path):
    p = tmp_path / "t.txt"
    data = gzip.open(GOLD_TXT, "rb").read()
    p.write_bytes(data[: len(data) // 2])
    with pytest.raises(RuntimeError):
        host.mesh_read(str(p))
    with pytest.raises(RuntimeError):
        host.mesh_read(str(tmp_path / "missing.txt"))


def test_voxelizer_rejects_empty_and_malformed(tmp_path):
    p = tmp_path / "e.off"
    p.write_text("OFF\n0 0 0\n")
    with pytest.raises(RuntimeError, match="empty"):
        host.voxelize_off(str(p), 1.0, 1.0, OUTSIDE)
    p.write_text("OFF\n3 1 0\n0 0 0\n1 0 0\n")
    with pytest.raises(RuntimeError, match="unreadable"):
        host.voxelize_off(str(p), 1.0, 1.0, OUTSIDE)


def test_synthetic_obstacle_scene_properties():
    v, f, objs = obstacle_scene()
    assert 19000 <= len(f) <= 21000
    assert edge_manifold(f)
    sv, sf = uv_ellipsoid((0, 0, 0), (24, 24, 24), 42, 60)
    # an inscribed polyhedron: volume slightly below the sphere's, within 1 %
    assert 0.99 < mesh_volume(sv, sf) / (4 / 3 * np.pi * 24 ** 3) < 1.0
    # obstacles inside the box, separated
    for k, (_, c, r) in enumerate(objs):
        assert np.all(c - r.max() > 0) and np.all(c + r.max() < np.array([512, 256, 256]))
        for _, c2, r2 in objs[k + 1:]:
            assert np.linalg.norm(c - c2) > r.max() + r2.max()


def test_voxelizer_against_oracle_and_analytic_shapes(tmp_path):
    """A small scene (box + sphere + ellipsoid): product == brute-force oracle
    everywhere, and == the analytic inside test away from the surfaces."""
    from lb_amd.geometry import box, merge

    parts = [box((0.0, 0.0, 0.0), (48.0, 32.0, 32.0)), uv_ellipsoid((14.0, 15.0, 17.0), (7.0, 7.0, 7.0), 24, 32),
             uv_ellipsoid((33.0, 16.0, 14.0), (6.0, 9.0, 5.0), 24, 32)]
    v, f = merge(parts)
    p = str(tmp_path / "scene.off")
    write_off(p, v, f)
    out = (-3.1416, -2.7183, -1.4142)
    g, mn, mx = host.voxelize_off(p, 1.0, 1.0, out)
    dims, go = oracle_voxelize(p, 1.0, 1.0, out)
    assert np.array_equal(g, go)
    nz, ny, nx = g.shape
    z, y, x = np.meshgrid(np.arange(nz) - 0.5, np.arange(ny) - 0.5, np.arange(nx) - 0.5, indexing="ij")
    s1 = ((x - 14) ** 2 + (y - 15) ** 2 + (z - 17) ** 2) / 49.0
    s2 = ((x - 33) / 6.0) ** 2 + ((y - 16) / 9.0) ** 2 + ((z - 14) / 5.0) ** 2
    inbox = (x > 0) & (x < 48) & (y > 0) & (y < 32) & (z > 0) & (z < 32)
    analytic = inbox & (s1 > 1) & (s2 > 1)
    far = (np.abs(np.sqrt(s1) - 1) > 0.15) & (np.abs(np.sqrt(s2) - 1) > 0.25) & \
          (np.minimum.reduce([np.abs(x), np.abs(x - 48), np.abs(y), np.abs(y - 32), np.abs(z), np.abs(z - 32)]) > 0.25)
    assert np.array_equal(g[far].astype(bool), analytic[far])
    assert far.mean() > 0.8


def _mesh_compare(grid, setup):
    """Run the same boundary meshing on the product mesher and the oracle
    mesher; compare flags and boundary lists."""
    m = host.Mesher(grid, 1.0, 1.0, (0.0, 0.0, 0.0))
    m.wall()
    of, walls = mo.meshing_boundary_wall(grid)
    obc = []
    for kind, args in setup:
        if kind == "inlet":
            p1, p2, p3, cid, nrm, vel = args
            assert m.inlet(p1, p2, p3, cid, nrm, vel, abi.BC_INLET_MOVING_WALL) == 0
            mask = mo.cuboid_mask(grid.shape, p1, p2, p3, lookup_dtype=np.float32)
            obc += [dict(kind=abi.BC_INLET_MOVING_WALL, **d) for d in mo.meshing_inlet_uniform(of, mask, cid, nrm, vel)]
        elif kind == "outlet":
            p1, p2, p3, cid, nrm, dens = args
            m.outlet(p1, p2, p3, cid, nrm, dens)
            mask = mo.cuboid_mask(grid.shape, p1, p2, p3)
            obc += [dict(kind=abi.BC_OUTLET_INTERPOLATED, **d) for d in mo.meshing_outlet(of, mask, cid, nrm, dens)]
        elif kind == "periodic":
            p1, p2, p3, cid, nrm, pi = args
            m.periodic(p1, p2, p3, cid, nrm, pi)
            mask = mo.cuboid_mask(grid.shape, p1, p2, p3)
            obc += [dict(kind=abi.BC_PERIODIC_GHOST, **d) for d in mo.meshing_periodic(of, mask, cid, nrm, pi)]
    walls = mo.compile_wall(of, walls)
    flags, bc = m.compile()
    assert np.array_equal(flags, of)
    pw = [(b.x, b.y, b.z, b.regular) for b in bc if b.kind == abi.BC_WALL]
    assert pw == [tuple(w) for w in walls]
    # the product lists inlets, then periodic, then outlets (compile order)
    order = {abi.BC_INLET_MOVING_WALL: 0, abi.BC_PERIODIC_GHOST: 1, abi.BC_OUTLET_INTERPOLATED: 2}
    pn = [(b.kind, b.x, b.y, b.z, b.regular) for b in bc if b.kind != abi.BC_WALL]
    on = sorted([(d["kind"], d["x"], d["y"], d["z"], d["regular"]) for d in obc],
                key=lambda t: order[t[0]])
    assert pn == on
    return flags, bc


def test_mesher_matches_oracle_cavity():
    n = 10
    g = np.zeros((n + 2, n + 2, n + 2), np.int32)
    g[1:-1, 1:-1, 1:-1] = 1
    flags, bc = _mesh_compare(g, [("inlet", ((-1.0, n - 0.75, -1.0), (n + 1.0, n - 0.25, n + 1.0),
                                              (-1.0, n - 0.75, n + 1.0), -1, (0.0, 1.0, 0.0), (0.05, 0.0, 0.0)))])
    # closed forms: face cells 1 + 9, edges 1 + 15, corners 1 + 19
    assert flags[1, 5, 5] == 10 and flags[1, 1, 5] == 16 and flags[1, 1, 1] == 20 and flags[5, 5, 5] == 1
    assert np.all(flags[1:-1, n, 1:-1] == -1)


def test_mesher_matches_oracle_voxelized_step():
    """Boundary meshing of the reference's own backward-step grid with an
    inlet, an outlet and two periodic cuboids."""
    dims, gold = golden_grid()
    setup = [
        # cell centres are at 4 i - 2 (resolution 0.25, one halo layer): the first
        # fluid layer sits at 2, the last one along x at 1498
        ("periodic", ((-1.0, 1.0, -1.0), (1600.0, 3.0, 300.0), (-1.0, 1.0, 300.0), -3, (0.0, -1.0, 0.0), 50)),
        ("inlet", ((1.0, -5.0, -5.0), (3.0, 250.0, 250.0), (1.0, 250.0, -5.0), -1, (-1.0, 0.0, 0.0),
                   (0.04, 0.0, 0.0))),
        ("outlet", ((1497.0, -5.0, -5.0), (1499.0, 250.0, 250.0), (1497.0, 250.0, -5.0), -2, (1.0, 0.0, 0.0), 1.0)),
    ]
    # the reference mesher works in physical coordinates: resolution 0.25, additional 1
    m = host.Mesher(gold, 0.25, 1.0, (0.0, 0.0, 0.0))
    m.wall()
    of, walls = mo.meshing_boundary_wall(gold)
    for kind, args in setup:
        p1, p2, p3, cid, nrm, last = args
        if kind == "inlet":
            assert m.inlet(p1, p2, p3, cid, nrm, last) == 0
            mo.meshing_inlet_uniform(of, mo.cuboid_mask(gold.shape, p1, p2, p3, 0.25, 1.0, lookup_dtype=np.float32),
                                     cid, nrm, last)
        elif kind == "outlet":
            m.outlet(p1, p2, p3, cid, nrm, last)
            mo.meshing_outlet(of, mo.cuboid_mask(gold.shape, p1, p2, p3, 0.25, 1.0), cid, nrm, last)
        else:
            m.periodic(p1, p2, p3, cid, nrm, last)
            mo.meshing_periodic(of, mo.cuboid_mask(gold.shape, p1, p2, p3, 0.25, 1.0), cid, nrm, last)
    flags, bc = m.compile()
    assert np.array_equal(flags, of)
    assert (flags == -1).sum() > 0 and (flags == -2).sum() > 0 and (flags == -3).sum() > 0
    assert sum(b.kind == abi.BC_WALL for b in bc) == len(mo.compile_wall(of, walls))


def test_inlet_with_equal_velocity_components_is_rejected():
    g = np.zeros((6, 6, 6), np.int32)
    g[1:-1, 1:-1, 1:-1] = 1
    m = host.Mesher(g)
    m.wall()
    assert m.inlet((0, 0, 0), (9, 9, 9), (0, 9, 9), -1, (1, 0, 0), (0.0, 0.0, 0.0)) == 1


def test_nondimensionalize_reference_arithmetic():
    """Unit conversion of NonDimensiolnaliseFactorsVelocity in float, with its
    tau = nu_lb / 3 + 1/2."""
    L = host.load_host()
    import ctypes

    h = ctypes.c_void_p()
    assert L.lbmh_case_create(ctypes.byref(h)) == 0
    g = np.ones((4, 4, 100), np.int32)
    dims = (ctypes.c_int32 * 3)(100, 4, 4)
    assert L.lbmh_case_set_grid(h, dims, g.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 100.0, 0.0,
                                (ctypes.c_float * 3)(0, 0, 0), (ctypes.c_float * 3)(1, 0.04, 0.04)) == 0
    for k, v in (("u_guess_fyz", 1.0), ("U_lb", 0.1), ("ny_fyz", 1e-3), ("rho_fyz", 1000.0)):
        assert L.lbmh_case_set(h, k.encode(), v) == 0
    assert L.lbmh_case_convert_to_lattice(h) == 0

    def get(k):
        v = ctypes.c_double()
        assert L.lbmh_case_get(h, k.encode(), ctypes.byref(v)) == 0
        return v.value

    f = np.float32
    Cl = f(1.0) / f(100)
    Cu = f(1.0) / f(0.1)
    Ct = Cl / Cu
    ny = f(1e-3) * Ct / Cl / Cl
    assert get("Cl") == Cl and get("Cu") == Cu and get("Ct") == Ct and get("ny") == ny
    assert get("tau") == ny / f(3) + f(0.5)
    assert get("omega") == f(1) / (ny / f(3) + f(0.5))
    assert abs(get("Re") - 1000.0) < 1e-3
    assert L.lbmh_case_set(h, b"no_such_field", 1.0) == 1
    L.lbmh_case_destroy(h)


def _vtk_fields_text(dims, rho, u, pressure=None):
    nx, ny, nz = dims
    lines = ["# vtk DataFile Version 2.0", "LBE variables", "ASCII", "DATASET STRUCTURED_POINTS",
             f"DIMENSIONS {nx} {ny} {nz}", "ASPECT_RATIO 1 1 1", "ORIGIN 0 0 0", f"POINT_DATA {nx * ny * nz}",
             "SCALARS rho double 1", "LOOKUP_TABLE default"]
    lines += [repr(float(v)) for v in rho.ravel()]
    lines += ["SCALARS pressure double 1", "LOOKUP_TABLE default"]
    lines += [repr(float(v)) for v in (pressure if pressure is not None else rho / 3).ravel()]
    lines += ["VECTORS U double"]
    lines += [f"{float(a)!r} {float(b)!r} {float(c)!r}" for a, b, c in zip(*(x.ravel() for x in u))]
    return "\n".join(lines) + "\n"


def test_vtk_initial_fields_reader(tmp_path):
    """InitializationEquilibriumVariables' reader: rho after the first
    LOOKUP_TABLE, u after VECTORS U double (the pressure block between is
    skipped), x fastest, every value rounded to float; short files fail."""
    import ctypes
    dims = (5, 4, 3)
    rng = np.random.default_rng(3)
    rho = 1 + 0.01 * rng.standard_normal(dims[::-1])
    u = [0.03 * rng.standard_normal(dims[::-1]) for _ in range(3)]
    p = tmp_path / "init.vtk"
    p.write_text(_vtk_fields_text(dims, rho, u))
    L = host.load_host()
    d = (ctypes.c_int32 * 3)(*dims)
    out = [np.empty(dims[::-1]) for _ in range(4)]
    assert L.lbmh_read_vtk_fields(str(p).encode(), d, *[abi.ptr(a) for a in out]) == 0
    for got, want in zip(out, [rho] + u):
        assert np.array_equal(got, want.astype(np.float32).astype(np.float64))
    text = p.read_text().splitlines()
    short_u = tmp_path / "short_u.vtk"
    short_u.write_text("\n".join(text[:-2]) + "\n")
    assert L.lbmh_read_vtk_fields(str(short_u).encode(), d, *[abi.ptr(a) for a in out]) == 3
    short_r = tmp_path / "short_r.vtk"
    short_r.write_text("\n".join(text[:15]) + "\n")
    assert L.lbmh_read_vtk_fields(str(short_r).encode(), d, *[abi.ptr(a) for a in out]) == 2
    assert L.lbmh_read_vtk_fields(str(tmp_path / "none.vtk").encode(), d, *[abi.ptr(a) for a in out]) == 1


# --- parabolic inlets (geometryMesherBoundary.h:111-274) -------------------

def _pipe_grid(nx=6, ny=17, nz=17):
    """A duct along x: solid rows on the y and z faces, fluid inside."""
    g = np.zeros((nz, ny, nx), np.int32)
    g[1:-1, 1:-1, :] = 1
    return g


def _inlet_lists(m, of, obc):
    flags, bc = m.compile()
    assert np.array_equal(flags, of)
    got = [(b.x, b.y, b.z, b.ux, b.uy, b.uz) for b in bc if b.kind == abi.BC_INLET_EQUILIBRIUM]
    want = [(d["x"], d["y"], d["z"], *[float(np.float32(v)) for v in d["velocity"]]) for d in obc]
    return got, want, bc


def test_paraboloid_circle_inlet_matches_restatement_and_closed_form():
    g = _pipe_grid()
    nz, ny, nx = g.shape
    # additional 1: cell j has centre j - 0.5; the duct's interior spans y, z in [0, 15];
    # the centre lies in the plane of the inlet cells' centres (the distance is 3-D)
    centre, R, U = (-0.5, 7.5, 7.5), 7.0, 0.04
    p1, p2, p3 = (-0.9, -1.0, -1.0), (-0.1, 20.0, 20.0), (-0.1, 20.0, -1.0)
    m = host.Mesher(g, 1.0, 1.0, (0.0, 0.0, 0.0))
    m.wall()
    of, _ = mo.meshing_boundary_wall(g)
    assert m.inlet_paraboloid_circle(p1, p2, p3, -1, (-1.0, 0.0, 0.0), centre, R, U) == 0
    obc = mo.meshing_inlet_paraboloid_circle(of, [], p1, p2, p3, -1, (-1.0, 0.0, 0.0), centre, R, U)
    got, want, bc = _inlet_lists(m, of, obc)
    assert got == want and len(got) == (ny - 2) * (nz - 2)
    # closed form: u_x = 3 U (1 - r^2 / R^2) inside R, 0 outside, into the domain (+x)
    for (x, y, z, ux, uy, uz) in got:
        r = math.hypot(y - 0.5 - centre[1], z - 0.5 - centre[2])
        expect = 3 * U * (1 - r * r / (R * R)) if r <= R else 0.0
        assert abs(ux - expect) < 1e-7 and uy == 0.0 and uz == 0.0
    # the cell at the centre carries the peak 3 U: the profile's mean over the
    # disc is 1.5 U (a pipe's Poiseuille peak would be 2 U)
    assert max(t[3] for t in got) == pytest.approx(3 * U, rel=1e-6)


def test_paraboloid_rectangle_inlet_matches_restatement_and_closed_form():
    g = _pipe_grid(nx=6, ny=13, nz=21)
    centre, dims, U = (0.0, 5.5, 9.5), (3.0, 11.0, 19.0), 0.03
    p1, p2, p3 = (-0.9, -1.0, -1.0), (-0.1, 30.0, 30.0), (-0.1, 30.0, -1.0)
    m = host.Mesher(g, 1.0, 1.0, (0.0, 0.0, 0.0))
    m.wall()
    of, _ = mo.meshing_boundary_wall(g)
    dims_after = m.inlet_paraboloid_rectangle(p1, p2, p3, -1, (-1.0, 0.0, 0.0), centre, dims, U)
    obc, dims_o = mo.meshing_inlet_paraboloid_rectangle(of, [], p1, p2, p3, -1, (-1.0, 0.0, 0.0), centre, dims, U)
    got, want, _ = _inlet_lists(m, of, obc)
    assert got == want and dims_after == dims_o == (1.0, 11.0, 19.0)
    # closed form: 2 U * prod 4 d (L - d) / L^2 over the two cross axes, d from the low corner
    for (x, y, z, ux, uy, uz) in got:
        dy, dz = (y - 0.5) - (centre[1] - dims[1] / 2), (z - 0.5) - (centre[2] - dims[2] / 2)
        expect = 2 * U * (4 * dy * (dims[1] - dy) / dims[1] ** 2) * (4 * dz * (dims[2] - dz) / dims[2] ** 2)
        assert abs(ux - max(expect, 0.0)) < 1e-7
    # the cell at the centre carries the peak 2 U
    assert max(t[3] for t in got) == pytest.approx(2 * U, rel=1e-6)


def test_paraboloid_overwrites_earlier_inlets_inside_its_cuboid():
    """The reference loops over every inlet entry, not only the new ones: an
    earlier uniform inlet inside the cuboid takes the parabolic velocity."""
    g = _pipe_grid()
    p1, p2, p3 = (-0.9, -1.0, -1.0), (-0.1, 20.0, 20.0), (-0.1, 20.0, -1.0)
    m = host.Mesher(g, 1.0, 1.0, (0.0, 0.0, 0.0))
    m.wall()
    of, _ = mo.meshing_boundary_wall(g)
    assert m.inlet(p1, p2, p3, -1, (-1.0, 0.0, 0.0), (0.01, 0.0, 0.0), abi.BC_INLET_EQUILIBRIUM) == 0
    mask = mo.cuboid_mask(g.shape, p1, p2, p3, lookup_dtype=np.float32)
    obc = mo.meshing_inlet_uniform(of, mask, -1, (-1.0, 0.0, 0.0), (0.01, 0.0, 0.0))
    m.inlet_paraboloid_circle(p1, p2, p3, -4, (-1.0, 0.0, 0.0), (-0.5, 7.5, 7.5), 7.0, 0.04)
    obc = mo.meshing_inlet_paraboloid_circle(of, obc, p1, p2, p3, -4, (-1.0, 0.0, 0.0), (-0.5, 7.5, 7.5), 7.0, 0.04)
    got, want, _ = _inlet_lists(m, of, obc)
    assert got == want
    for (x, y, z, ux, uy, uz) in got:
        inside = math.hypot(y - 0.5 - 7.5, z - 0.5 - 7.5) <= 7.0
        assert (ux != np.float32(0.01)) == inside      # outside the radius the old velocity stays
    assert np.all(of[:, :, 0][g[:, :, 0] == 1] == -1)    # the cells keep the first inlet's id
