"""Restatement of the reference's boundary mesher (flag-grid post-processing).

TEST INFRASTRUCTURE ONLY: the checker for the product mesher
(lb_amd/csrc/host/geometry_mesher_boundary.cpp) and a generator of test
geometries.  Follows 3D/simulation/src/geometry/geometryMesherBoundary.h.

Arrays are indexed [z, y, x] (x fastest in memory), like the reference's
meshFluidHost with index_sequence<2,1,0>.
"""
from __future__ import annotations

import math

import numpy as np

# lbm_bc_node kinds (include/lbm_core.h)
BC_WALL, BC_INLET_MOVING_WALL, BC_INLET_EQUILIBRIUM, BC_OUTLET_INTERPOLATED, BC_PERIODIC_GHOST = range(5)
BC_OUTLET_EQUILIBRIUM, BC_OUTLET_NEIGHBOUR_EQ, BC_OUTLET_NEIGHBOUR_EQ_RF, BC_INLET_ZOU_HE = 5, 6, 7, 8


def meshing_boundary_wall(grid: np.ndarray):
    """geometryMesherBoundary::meshingBoundaryWall (:439-484).

    Every fluid cell (value > 0) gets +1 per solid cell (value 0) in its
    in-bounds 3x3x3 neighbourhood.  Cells whose count exceeds 2 enter the wall
    list in k, j, i loop order; regular = (count <= 10).
    Returns (flags, wall_list[(x, y, z, regular)]).
    """
    g = np.asarray(grid, dtype=np.int32)
    solid = (g == 0).astype(np.int32)
    nz, ny, nx = g.shape
    pad = np.zeros((nz + 2, ny + 2, nx + 2), np.int32)
    pad[1:-1, 1:-1, 1:-1] = solid
    cnt = np.zeros_like(g)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                cnt += pad[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
    flags = np.where(g > 0, g + cnt, g).astype(np.int32)
    zz, yy, xx = np.nonzero(flags > 2)
    order = np.lexsort((xx, yy, zz))  # k outer, then j, then i
    walls = [(int(xx[o]), int(yy[o]), int(zz[o]), int(flags[zz[o], yy[o], xx[o]] <= 10)) for o in order]
    return flags, walls


def cell_centre(i, j, k, resolution, additional, bbmin):
    """Physical query point of a cell (geometryMesherBoundary.h:74-76)."""
    return ((i + 0.5 - additional) / resolution + bbmin[0],
            (j + 0.5 - additional) / resolution + bbmin[1],
            (k + 0.5 - additional) / resolution + bbmin[2])


def cuboid_mask(shape, p1, p2, p3, resolution=1.0, additional=1.0, bbmin=(0.0, 0.0, 0.0), lookup_dtype=np.float64):
    """geometryObjectCuboid::isInside (:21-38) evaluated at every cell centre.

    lookup_dtype is the type the query coordinates are held in before the
    test (RealType lookx in the inlet mesher, double in the others)."""
    nz, ny, nx = shape
    lo = [min(p1[a], p2[a], p3[a]) for a in range(3)]
    hi = [max(p1[a], p2[a], p3[a]) for a in range(3)]
    i = np.arange(nx, dtype=np.float64)
    j = np.arange(ny, dtype=np.float64)
    k = np.arange(nz, dtype=np.float64)
    lx = ((i + 0.5 - additional) / resolution + bbmin[0]).astype(lookup_dtype).astype(np.float64)
    ly = ((j + 0.5 - additional) / resolution + bbmin[1]).astype(lookup_dtype).astype(np.float64)
    lz = ((k + 0.5 - additional) / resolution + bbmin[2]).astype(lookup_dtype).astype(np.float64)
    mx = (lx >= lo[0]) & (lx <= hi[0])
    my = (ly >= lo[1]) & (ly <= hi[1])
    mz = (lz >= lo[2]) & (lz <= hi[2])
    return mz[:, None, None] & my[None, :, None] & mx[None, None, :]


def _scan(mask):
    zz, yy, xx = np.nonzero(mask)
    order = np.lexsort((xx, yy, zz))
    return [(int(xx[o]), int(yy[o]), int(zz[o])) for o in order]


def meshing_inlet_uniform(flags, mask, cuboid_id, normal, velocity):
    """meshingBoundaryConditionInletUniform (:47-108): fluid cells (flag > 0)
    inside the cuboid become cuboid_id; regular = (flag <= 10)."""
    nodes = []
    for (x, y, z) in _scan(mask & (flags > 0)):
        reg = int(flags[z, y, x] <= 10)
        flags[z, y, x] = cuboid_id
        nodes.append(dict(x=x, y=y, z=z, normal=tuple(normal), velocity=tuple(velocity), regular=reg))
    return nodes


def _inside(p1, p2, p3, q):
    lo = [min(p1[a], p2[a], p3[a]) for a in range(3)]
    hi = [max(p1[a], p2[a], p3[a]) for a in range(3)]
    return all(lo[a] <= q[a] <= hi[a] for a in range(3))


def _centre32(d, resolution, additional, bbmin):
    """Cell centre of an entry held in float (RealType lookx), as floats."""
    c = cell_centre(d["x"], d["y"], d["z"], resolution, additional, bbmin)
    return tuple(float(np.float32(v)) for v in c)


def _set_velocity(d, v, normal):
    f = np.float32
    d["velocity"] = tuple(float(f(-v) * f(n)) for n in normal)


def meshing_inlet_paraboloid_circle(flags, inlets, p1, p2, p3, cuboid_id, normal, center, radius, mean,
                                    resolution=1.0, additional=1.0, bbmin=(0.0, 0.0, 0.0)):
    """meshingBoundaryConditionInletParaboloidCircle (:111-178): the uniform
    rule with velocity 0 appends the cuboid's cells to `inlets`, then every
    entry whose float centre is in the cuboid and within `radius` of `center`
    gets v = 3 mean (1 - (r/R)^2) against the normal (r in double, held in
    float; the square in double)."""
    f = np.float32
    mask = cuboid_mask(flags.shape, p1, p2, p3, resolution, additional, bbmin, lookup_dtype=np.float32)
    inlets += meshing_inlet_uniform(flags, mask, cuboid_id, normal, (0.0, 0.0, 0.0))
    for d in inlets:
        q = _centre32(d, resolution, additional, bbmin)
        if not _inside(p1, p2, p3, q):
            continue
        dist = f(math.sqrt(math.pow(q[0] - center[0], 2) + math.pow(q[1] - center[1], 2) +
                           math.pow(q[2] - center[2], 2)))
        if dist <= f(radius):
            v = f(float(f(3) * f(mean)) * (1.0 - math.pow(float(dist / f(radius)), 2)))
            _set_velocity(d, max(v, f(0)), normal)
    return inlets


def meshing_inlet_paraboloid_rectangle(flags, inlets, p1, p2, p3, cuboid_id, normal, center, dims, mean,
                                       resolution=1.0, additional=1.0, bbmin=(0.0, 0.0, 0.0)):
    """meshingBoundaryConditionInletParaboloidRectangle (:180-274): a product of
    1-D parabolas d_i (L_i - d_i) over the axes across the normal, divided by
    prod (L_i / 2)^2 (float, the reference's operation order), times 2 mean.
    Returns the inlets and the extents, the one along the normal set to 1."""
    f = np.float32
    L = [f(v) for v in dims]
    mask = cuboid_mask(flags.shape, p1, p2, p3, resolution, additional, bbmin, lookup_dtype=np.float32)
    inlets += meshing_inlet_uniform(flags, mask, cuboid_id, normal, (0.0, 0.0, 0.0))
    for d in inlets:
        q = _centre32(d, resolution, additional, bbmin)
        if not _inside(p1, p2, p3, q):
            continue
        part = []
        for a in range(3):
            dist = f(q[a] - (center[a] - float(L[a] / f(2))))
            if normal[a] == 0:
                part.append(dist * (L[a] - dist))
            else:
                part.append(f(1))
                L[a] = f(1)
        const = L[0] * L[1] * L[2] * L[0] * L[1] * L[2] / f(16)
        v = f(2) * f(mean) * part[0] * part[1] * part[2] / const
        _set_velocity(d, max(v, f(0)), normal)
    return inlets, tuple(float(v) for v in L)


def meshing_outlet(flags, mask, cuboid_id, normal, density):
    """meshingBoundaryConditionOutlet (:276-325): non-solid cells inside the
    cuboid become cuboid_id; regular = (flag == 10)."""
    nodes = []
    for (x, y, z) in _scan(mask & (flags != 0)):
        reg = int(flags[z, y, x] == 10)
        flags[z, y, x] = cuboid_id
        nodes.append(dict(x=x, y=y, z=z, normal=tuple(normal), density=float(density), regular=reg))
    return nodes


def meshing_periodic(flags, mask, cuboid_id, normal, periodic_index):
    """meshingBoundaryConditionPeriodic (:379-437)."""
    nodes = []
    for (x, y, z) in _scan(mask & (flags != 0)):
        reg = int(flags[z, y, x] == 10)
        flags[z, y, x] = cuboid_id
        nodes.append(dict(x=x, y=y, z=z, normal=tuple(normal), per_idx=int(periodic_index), regular=reg))
    return nodes


def compile_wall(flags, walls):
    """compileBoundaryArrayWall (:521-561): drop entries on negative flags."""
    return [w for w in walls if flags[w[2], w[1], w[0]] >= 0]


def cavity(n: int, lid_velocity=(0.05, 0.0, 0.0)):
    """Lid-driven cavity: n^3 fluid cells inside a one-cell solid halo; the
    top fluid layer (y = n) is a moving-wall inlet with outward normal +y.

    Returns (flags, bc) with bc a list of dicts {kind, x, y, z, ...}.
    """
    g = np.zeros((n + 2, n + 2, n + 2), np.int32)
    g[1:-1, 1:-1, 1:-1] = 1
    flags, walls = meshing_boundary_wall(g)
    # lid cuboid: cell centres (j + 0.5 - 1) in [n - 1, n] selects j = n only
    mask = cuboid_mask(flags.shape, (-1.0, n - 0.75, -1.0), (n + 1.0, n - 0.25, n + 1.0), (-1.0, n - 0.75, n + 1.0),
                       lookup_dtype=np.float32)
    vel = tuple(float(np.float32(v)) for v in lid_velocity)
    inl = meshing_inlet_uniform(flags, mask, -1, (0.0, 1.0, 0.0), vel)
    walls = compile_wall(flags, walls)
    bc = [dict(kind=BC_WALL, x=w[0], y=w[1], z=w[2], regular=w[3]) for w in walls]
    bc += [dict(kind=BC_INLET_MOVING_WALL, **d) for d in inl]
    return flags, bc


def channel(nx, ny, nz, u_in=0.04, rho_out=1.0, ghost_z=True, obstacle=True, inlet_kind=BC_INLET_EQUILIBRIUM,
            outlet_kind=BC_OUTLET_INTERPOLATED,
            walls=True):
    """Channel along x: solid rows at y = 0 and y = ny - 1, an equilibrium
    velocity inlet on x = 0 (outward normal -x), an interpolated density
    outlet on x = nx - 1 (+x) and, with ghost_z, ghost-layer periodicity
    along z (z = 0 copies from z = nz - 2, z = nz - 1 from z = 1, the pattern
    of the reference's main.cu:178-179).  Cells sit at i + 0.5 (additional 0).

    walls=False drops the solid rows (the caller then wraps y).
    Meshing order: wall, periodic, inlet, outlet, compile -- so the edge
    cells (x = nx - 1, z = 0 / nz - 1) are both outlet and ghost nodes and
    the ghost nodes at x = 0 have inlet cells as partners.
    Returns (flags, bc) with bc a list of dicts."""
    g = np.ones((nz, ny, nx), np.int32)
    if walls:
        g[:, 0, :] = 0
        g[:, ny - 1, :] = 0
    if obstacle:
        x0, y0 = nx // 3, ny // 2 - 1
        g[2:nz - 2, y0:y0 + 2, x0:x0 + 2] = 0
    flags, walls = meshing_boundary_wall(g)
    kw = dict(additional=0.0)
    shape = flags.shape
    bc = []
    per = []
    if ghost_z:
        m0 = cuboid_mask(shape, (-1.0, -1.0, 0.2), (nx + 1.0, ny + 1.0, 0.8), (-1.0, ny + 1.0, 0.8), **kw)
        per += meshing_periodic(flags, m0, -3, (0.0, 0.0, -1.0), nz - 2)
        m1 = cuboid_mask(shape, (-1.0, -1.0, nz - 0.8), (nx + 1.0, ny + 1.0, nz - 0.2), (-1.0, ny + 1.0, nz - 0.2),
                         **kw)
        per += meshing_periodic(flags, m1, -3, (0.0, 0.0, 1.0), 1)
    mi = cuboid_mask(shape, (0.2, -1.0, -1.0), (0.8, ny + 1.0, nz + 1.0), (0.8, ny + 1.0, -1.0),
                     lookup_dtype=np.float32, **kw)
    vel = (float(np.float32(u_in)), 0.0, 0.0)
    inl = meshing_inlet_uniform(flags, mi, -1, (-1.0, 0.0, 0.0), vel)
    mo = cuboid_mask(shape, (nx - 0.8, -1.0, -1.0), (nx - 0.2, ny + 1.0, nz + 1.0), (nx - 0.2, ny + 1.0, -1.0), **kw)
    out = meshing_outlet(flags, mo, -2, (1.0, 0.0, 0.0), rho_out)
    walls = compile_wall(flags, walls)
    bc += [dict(kind=BC_WALL, x=w[0], y=w[1], z=w[2], regular=w[3]) for w in walls]
    bc += [dict(kind=inlet_kind, **d) for d in inl]
    bc += [dict(kind=outlet_kind, **d) for d in out]
    bc += [dict(kind=BC_PERIODIC_GHOST, **d) for d in per]
    return flags, bc


def ghost_box(m):
    """m^3 fluid cells with ghost-layer periodicity on all six faces
    (x faces first, then y, then z; face cuboids include edges and corners,
    so those cells appear in two or three periodic entries)."""
    g = np.ones((m, m, m), np.int32)
    flags, _ = meshing_boundary_wall(g)
    shape = flags.shape
    per = []
    big = m + 1.0
    for axis in range(3):
        for side, (lo, hi, idx) in enumerate(((0.2, 0.8, m - 2), (m - 0.8, m - 0.2, 1))):
            p1 = [-1.0, -1.0, -1.0]
            p2 = [big, big, big]
            p1[axis], p2[axis] = lo, hi
            p3 = list(p2)
            p3[0] = p1[0]
            nrm = [0.0, 0.0, 0.0]
            nrm[axis] = -1.0 if side == 0 else 1.0
            mask = cuboid_mask(shape, tuple(p1), tuple(p2), tuple(p3), additional=0.0)
            per += meshing_periodic(flags, mask, -3, tuple(nrm), idx)
    return flags, [dict(kind=BC_PERIODIC_GHOST, **d) for d in per]
