"""Restatement of the reference's boundary mesher (flag-grid post-processing).

TEST INFRASTRUCTURE ONLY: the checker for the product mesher
(lb_amd/csrc/host/geometry_mesher_boundary.cpp) and a generator of test
geometries.  Follows 3D/simulation/src/geometry/geometryMesherBoundary.h.

Arrays are indexed [z, y, x] (x fastest in memory), like the reference's
meshFluidHost with index_sequence<2,1,0>.
"""
from __future__ import annotations

import numpy as np

# lbm_bc_node kinds (include/lbm_core.h)
BC_WALL, BC_INLET_MOVING_WALL, BC_INLET_EQUILIBRIUM, BC_OUTLET_INTERPOLATED, BC_PERIODIC_GHOST = range(5)


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
