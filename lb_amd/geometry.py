"""Synthetic OFF surfaces for the obstacle configuration (host side, numpy).

obstacle_scene(): the domain box [0, 512] x [0, 256] x [0, 256] (outward
oriented, so cells inside the box are fluid) plus a sphere of radius 24 at a
seeded centre and three seeded, mutually non-overlapping ellipsoids; every
obstacle is a closed triangulated surface (every edge shared by exactly two
faces), so cells inside an obstacle see an even number of crossings and are
solid.  About 20k triangles in total.
"""
from __future__ import annotations

import numpy as np


def uv_ellipsoid(center, radii, n_theta, n_phi):
    """Closed triangulated ellipsoid: poles plus n_theta-1 rings of n_phi vertices."""
    cx, cy, cz = center
    a, b, c = radii
    verts = [(cx, cy, cz + c)]
    for i in range(1, n_theta):
        th = np.pi * i / n_theta
        for j in range(n_phi):
            ph = 2 * np.pi * j / n_phi
            verts.append((cx + a * np.sin(th) * np.cos(ph), cy + b * np.sin(th) * np.sin(ph), cz + c * np.cos(th)))
    verts.append((cx, cy, cz - c))
    south = len(verts) - 1
    tris = []
    ring = lambda i, j: 1 + (i - 1) * n_phi + (j % n_phi)  # noqa: E731
    for j in range(n_phi):
        tris.append((0, ring(1, j), ring(1, j + 1)))
    for i in range(1, n_theta - 1):
        for j in range(n_phi):
            a0, a1 = ring(i, j), ring(i, j + 1)
            b0, b1 = ring(i + 1, j), ring(i + 1, j + 1)
            tris.append((a0, b0, b1))
            tris.append((a0, b1, a1))
    for j in range(n_phi):
        tris.append((south, ring(n_theta - 1, j + 1), ring(n_theta - 1, j)))
    return np.array(verts, dtype=np.float64), np.array(tris, dtype=np.int64)


def box(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0),
                  (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)], dtype=np.float64)
    f = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4),
         (2, 3, 7), (2, 7, 6), (1, 2, 6), (1, 6, 5), (0, 4, 7), (0, 7, 3)]
    return v, np.array(f, dtype=np.int64)


def merge(parts):
    vs, fs, off = [], [], 0
    for v, f in parts:
        vs.append(v)
        fs.append(f + off)
        off += len(v)
    return np.concatenate(vs), np.concatenate(fs)


def obstacle_scene(seed=20240603, domain=(512.0, 256.0, 256.0), sphere_radius=24.0, n_theta=42, n_phi=60,
                   ellipsoid_radii=((10, 8, 8), (28, 20, 20)), margin=4.0):
    """Returns (vertices, triangles, objects) where objects lists the analytic
    shapes: ('sphere'|'ellipsoid', centre, radii).  The defaults are the
    obstacle workload; smaller domains take smaller radii and margins."""
    rng = np.random.default_rng(seed)
    L = np.array(domain)
    objs = []
    c = np.array([rng.uniform(0.2, 0.3) * L[0], rng.uniform(0.4, 0.6) * L[1], rng.uniform(0.4, 0.6) * L[2]])
    objs.append(("sphere", c, np.array([sphere_radius] * 3)))
    tries = 0
    while len(objs) < 4 and tries < 1000:
        tries += 1
        r = rng.uniform(ellipsoid_radii[0], ellipsoid_radii[1])
        cen = rng.uniform([0.35 * L[0], 0.15 * L[1], 0.15 * L[2]], [0.85 * L[0], 0.85 * L[1], 0.85 * L[2]])
        ok = all(np.linalg.norm(cen - o[1]) > r.max() + o[2].max() + margin for o in objs)
        ok = ok and np.all(cen - r.max() > margin) and np.all(cen + r.max() < L - margin)
        if ok:
            objs.append(("ellipsoid", cen, r))
    parts = [box((0.0, 0.0, 0.0), tuple(L))]
    for kind, cen, r in objs:
        parts.append(uv_ellipsoid(tuple(cen), tuple(r), n_theta, n_phi))
    v, f = merge(parts)
    return v, f, objs


def write_off(path, verts, tris):
    with open(path, "w") as fh:
        fh.write("OFF\n")
        fh.write(f"{len(verts)} {len(tris)} 0\n")
        for v in verts:
            fh.write(f"{v[0]:.17g} {v[1]:.17g} {v[2]:.17g}\n")
        for t in tris:
            fh.write(f"3 {t[0]} {t[1]} {t[2]}\n")


def edge_manifold(tris):
    """True when every undirected edge is shared by exactly two faces."""
    from collections import Counter

    cnt = Counter()
    for a, b, c in tris:
        for e in ((a, b), (b, c), (c, a)):
            cnt[tuple(sorted(e))] += 1
    return all(v == 2 for v in cnt.values())


def mesh_volume(verts, tris):
    """Signed volume by the divergence theorem (positive for outward faces)."""
    a, b, c = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    return float(np.sum(np.einsum("ij,ij->i", a, np.cross(b, c))) / 6.0)
