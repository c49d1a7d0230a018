"""Seeded synthetic meshes: jittered split grids on the unit square / cube.

* 2D: an ``n x n`` grid of the unit square, each cell split into two
  triangles, along the diagonal (i,j)-(i+1,j+1) (``diagonal=0``) or along
  (i+1,j)-(i,j+1) (``diagonal=1``).  ``2 n^2`` triangles.
* 3D: an ``n^3`` grid of the unit cube, each hexahedron split into the six
  Kuhn tetrahedra around its main diagonal.  ``6 n^3`` tetrahedra.

Every interior vertex is displaced by an independent uniform draw in
``[-alpha h, alpha h]`` per coordinate (``h = 1/n``); boundary vertices stay
fixed, so the meshed domain is exactly the unit square / cube.  Draws come
from a counter-based hash of ``(seed, mesh_id, draw, vertex, coordinate)``,
so the same mesh is produced on any machine in any order.

Inversion rejection: an element whose signed measure falls below 10% of its
unjittered signed measure would make the mesh overlap itself.  Every interior
vertex of such an element is re-drawn (draw index 1, 2, ...) for up to
``max_redraw`` rounds; vertices of elements still bad after that are reset to
their grid position, round after round, until no bad element is left.
"""
from __future__ import annotations

import math

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, mesh_id, draw, counters):
    """Uniform doubles in [0, 1) keyed by (seed, mesh_id, draw, counter)."""
    k = _splitmix64(np.uint64(seed))
    k = _splitmix64(k ^ np.uint64(mesh_id))
    k = _splitmix64(k ^ (np.uint64(draw) << np.uint64(48)))
    h = _splitmix64(k ^ np.asarray(counters, dtype=np.uint64))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _grid_vertices(dim, n):
    t = np.arange(n + 1, dtype=np.float64) / float(n)
    if dim == 2:
        y, x = np.meshgrid(t, t, indexing="ij")
        return np.stack([x.ravel(), y.ravel()], axis=1)
    z, y, x = np.meshgrid(t, t, t, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)


def _interior_mask(dim, n):
    i = np.arange(n + 1)
    inner = (i > 0) & (i < n)
    if dim == 2:
        jj, ii = np.meshgrid(inner, inner, indexing="ij")
        return (jj & ii).ravel()
    kk, jj, ii = np.meshgrid(inner, inner, inner, indexing="ij")
    return (kk & jj & ii).ravel()


def _elements(dim, n, diagonal):
    if dim == 2:
        j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        v00 = (j * (n + 1) + i).ravel()
        v10, v01, v11 = v00 + 1, v00 + (n + 1), v00 + (n + 2)
        if diagonal == 0:
            tris = [(v00, v10, v11), (v00, v11, v01)]
        else:
            tris = [(v00, v10, v01), (v10, v11, v01)]
        out = np.empty((n * n, 2, 3), dtype=np.int64)
        for t, (a, b, c) in enumerate(tris):
            out[:, t, 0], out[:, t, 1], out[:, t, 2] = a, b, c
        return out.reshape(-1, 3).astype(np.int32)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    v000 = ((k * (n + 1) + j) * (n + 1) + i).ravel()
    step = [1, n + 1, (n + 1) * (n + 1)]
    v111 = v000 + step[0] + step[1] + step[2]
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    out = np.empty((n ** 3, 6, 4), dtype=np.int64)
    for t, (p0, p1, _) in enumerate(perms):
        out[:, t, 0] = v000
        out[:, t, 1] = v000 + step[p0]
        out[:, t, 2] = v000 + step[p0] + step[p1]
        out[:, t, 3] = v111
    return out.reshape(-1, 4).astype(np.int32)


def signed_measures(pos, enl):
    """Signed area (2D) or volume (3D) of every element."""
    dim = pos.shape[1]
    p = pos[enl]
    e = p[:, 1:, :] - p[:, :1, :]
    if dim == 2:
        return 0.5 * (e[:, 0, 0] * e[:, 1, 1] - e[:, 0, 1] * e[:, 1, 0])
    cx = e[:, 1, 1] * e[:, 2, 2] - e[:, 1, 2] * e[:, 2, 1]
    cy = e[:, 1, 2] * e[:, 2, 0] - e[:, 1, 0] * e[:, 2, 2]
    cz = e[:, 1, 0] * e[:, 2, 1] - e[:, 1, 1] * e[:, 2, 0]
    return (e[:, 0, 0] * cx + e[:, 0, 1] * cy + e[:, 0, 2] * cz) / 6.0


def split_grid(dim, n, seed, mesh_id=0, alpha=None, diagonal=0, max_redraw=16):
    """Jittered split grid: ``(positions float64[nn, dim], enlist int32[ne, dim+1])``."""
    if dim not in (2, 3):
        raise ValueError("Invalid dimension")
    if n < 1:
        raise ValueError("n must be >= 1")
    if alpha is None:
        alpha = 0.3 if dim == 2 else 0.25
    base = _grid_vertices(dim, n)
    enl = _elements(dim, n, diagonal)
    nn = base.shape[0]
    interior = _interior_mask(dim, n)
    h = 1.0 / float(n)
    counters = np.arange(nn * dim, dtype=np.uint64).reshape(nn, dim)
    draw = np.zeros(nn, dtype=np.int64)

    def displacement(vertices, draws):
        out = np.empty((vertices.size, dim), dtype=np.float64)
        for dval in np.unique(draws):
            sel = draws == dval
            u = uniform01(seed, mesh_id, int(dval), counters[vertices[sel]])
            out[sel] = (2.0 * u - 1.0) * (alpha * h)
        return out

    disp = np.zeros((nn, dim), dtype=np.float64)
    ids = np.nonzero(interior)[0]
    disp[ids] = displacement(ids, draw[ids])
    ref = signed_measures(base, enl)
    pos = base + disp
    for _ in range(max_redraw):
        bad = signed_measures(pos, enl) / ref < 0.1
        if not bad.any():
            return pos, enl
        verts = np.unique(enl[bad].ravel())
        verts = verts[interior[verts]]
        draw[verts] += 1
        disp[verts] = displacement(verts, draw[verts])
        pos = base + disp
    while True:
        bad = signed_measures(pos, enl) / ref < 0.1
        if not bad.any():
            return pos, enl
        verts = np.unique(enl[bad].ravel())
        disp[verts] = 0.0
        pos = base + disp


def cell_grid(dim, n, seed, mesh_id=0, alpha=0.2):
    """Grid of the unit square / cube in quadrilaterals (2D, ``n^2``) or
    hexahedra (3D, ``n^3``), nodes in the reference's element order (2D:
    counter-clockwise; 3D: the z-low face counter-clockwise, then the z-high
    face).  2D: interior vertices jittered by up to ``alpha h`` per coordinate
    (convex cells for alpha < 0.25).  3D: the grid lines of each axis are
    jittered instead, so every face stays planar and the hexahedra tile the
    cube exactly.  Returns ``(positions float64[nn, dim], enlist int32[ne,
    2^dim])``."""
    if dim not in (2, 3):
        raise ValueError("Invalid dimension")
    if n < 1:
        raise ValueError("n must be >= 1")
    h = 1.0 / float(n)
    base = _grid_vertices(dim, n)
    nn = base.shape[0]
    if dim == 2:
        interior = _interior_mask(dim, n)
        u = uniform01(seed, mesh_id, 0, np.arange(nn * dim, dtype=np.uint64)).reshape(nn, dim)
        pos = base + np.where(interior[:, None], (2.0 * u - 1.0) * (alpha * h), 0.0)
        i, j = (x.ravel() for x in np.meshgrid(np.arange(n), np.arange(n), indexing="ij"))
        v = lambda a, b: b * (n + 1) + a  # noqa: E731
        enl = np.stack([v(i, j), v(i + 1, j), v(i + 1, j + 1), v(i, j + 1)], axis=1)
    else:
        u = uniform01(seed, mesh_id, 0, np.arange(3 * (n + 1), dtype=np.uint64)).reshape(3, n + 1)
        t = np.arange(n + 1, dtype=np.float64) * h
        axes = [t + np.where((t > 0) & (t < 1), (2.0 * u[d] - 1.0) * (alpha * h), 0.0) for d in range(3)]
        z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        pos = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
        i, j, k = (x.ravel() for x in np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"))
        v = lambda a, b, c: (c * (n + 1) + b) * (n + 1) + a  # noqa: E731
        enl = np.stack([v(i, j, k), v(i + 1, j, k), v(i + 1, j + 1, k), v(i, j + 1, k),
                        v(i, j, k + 1), v(i + 1, j, k + 1), v(i + 1, j + 1, k + 1), v(i, j + 1, k + 1)],
                       axis=1)
    return pos, enl.astype(np.int32)


# Named workloads: (dim, n_a, n_b, seed, alpha)
CONFIGS = {
    "tri_64_vs_50": (2, 64, 50, 0, 0.3),
    "tri_1024_vs_1000": (2, 1024, 1000, 1, 0.3),
    "tet_64_vs_60": (3, 64, 60, 2, 0.25),
    "tet_96_vs_90": (3, 96, 90, 4, 0.25),
}


def config_meshes(name):
    """``(pos_a, enl_a, pos_b, enl_b)`` of a named workload.  Mesh A splits
    along one diagonal and mesh B along the other (2D); both are Kuhn grids
    in 3D."""
    dim, na, nb, seed, alpha = CONFIGS[name]
    pa, ea = split_grid(dim, na, seed, mesh_id=0, alpha=alpha, diagonal=0)
    pb, eb = split_grid(dim, nb, seed, mesh_id=1, alpha=alpha, diagonal=1)
    return pa, ea, pb, eb


def batched_pairs(count=4096, seed=3, emin=50, emax=50000):
    """Mesh pairs of uneven sizes for the batched workload: each pair is two
    jittered split grids (2D or 3D with equal odds) whose element counts are
    drawn log-uniformly in [emin, emax] independently per mesh.  Returns a list
    of ``(dim, pos_a, enl_a, pos_b, enl_b)``."""
    u = uniform01(seed, 1 << 20, 0, np.arange(3 * count, dtype=np.uint64)).reshape(count, 3)
    out = []
    for p in range(count):
        dim = 2 if u[p, 0] < 0.5 else 3
        meshes = []
        for m in range(2):
            e = math.exp(math.log(emin) + u[p, 1 + m] * (math.log(emax) - math.log(emin)))
            if dim == 2:
                n = max(1, int(round(math.sqrt(e / 2.0))))
            else:
                n = max(1, int(round((e / 6.0) ** (1.0 / 3.0))))
            meshes.append(split_grid(dim, n, seed, mesh_id=2 * p + m, diagonal=m))
        out.append((dim, meshes[0][0], meshes[0][1], meshes[1][0], meshes[1][1]))
    return out
