"""Synthetic meshes for the indirect par_loop workloads.

Each generator returns plain numpy arrays (int32 maps, float64 coordinates)
that are handed to ``op2.Map`` / ``op2.Dat`` exactly as a mesh library would.

* :func:`unit_square_tri` -- structured ``n x n`` unit square, every square
  split into two triangles along the (i, j) -> (i+1, j+1) diagonal, interior
  vertices jittered, optional random vertex renumbering.
* :func:`unit_cube_tet_p2` -- ``n^3`` unit cube, every hex split into the six
  Kuhn tetrahedra around its main diagonal, with the quadratic (P2) node
  numbering "vertices first, then edges ordered by their sorted vertex pair".

Random draws follow one fixed order per workload so that a seed reproduces
every array: jitter x, jitter y, u, vertex permutation, then anything else.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

__all__ = ["TriMesh", "TetP2Mesh", "unit_square_tri", "unit_cube_tet_p2",
           "batched_tri_meshes", "workload_rng", "TET_EDGES"]

# Local edge -> local vertex pair of a tetrahedron; P2 local node 4+k sits on
# edge k.  The element kernels and the HIP kernels use the same table.
TET_EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))


def workload_rng(config):
    """Seeded generator for workload ``config`` (1..5)."""
    return np.random.default_rng([2026, int(config)])


@dataclass
class TriMesh:
    n: int
    coords: np.ndarray          # (nverts, 2) float64
    cell_vertex: np.ndarray     # (ncells, 3) int32
    u: Optional[np.ndarray]     # (nverts,) float64 or None
    perm: Optional[np.ndarray] = None   # old vertex id -> new vertex id

    @property
    def ncells(self):
        return self.cell_vertex.shape[0]

    @property
    def nverts(self):
        return self.coords.shape[0]


@dataclass
class TetP2Mesh:
    n: int
    coords: np.ndarray          # (nverts, 3) float64
    cell_vertex: np.ndarray     # (ncells, 4) int32
    cell_node: np.ndarray       # (ncells, 10) int32, P2 nodes
    nedges: int
    u: Optional[np.ndarray] = None      # (nnodes, 3) float64
    extra: dict = field(default_factory=dict)

    @property
    def ncells(self):
        return self.cell_vertex.shape[0]

    @property
    def nverts(self):
        return self.coords.shape[0]

    @property
    def nnodes(self):
        return self.nverts + self.nedges


def unit_square_tri(n, *, rng=None, jitter=0.2, draw_u=True, permute=False):
    """``n x n`` triangulated unit square (2 n^2 cells, (n+1)^2 vertices).

    Vertex (i, j) has id ``j*(n+1) + i`` before permutation.  Square (i, j)
    yields cells ``(v00, v10, v11)`` and ``(v00, v11, v01)`` (counter-clockwise).
    Interior vertices move by U(-jitter*h, jitter*h) in x then in y.  With
    ``permute`` the vertex ids are relabelled by ``rng.permutation``.
    """
    n = int(n)
    if n < 1:
        raise ValueError("n must be >= 1")
    n1 = n + 1
    h = 1.0 / n
    jj, ii = np.divmod(np.arange(n1 * n1, dtype=np.int64), n1)
    x = ii * h
    y = jj * h
    interior = (ii > 0) & (ii < n) & (jj > 0) & (jj < n)
    nint = int(interior.sum())
    if rng is not None and jitter:
        jx = rng.uniform(-jitter * h, jitter * h, size=nint)
        jy = rng.uniform(-jitter * h, jitter * h, size=nint)
        x = x.copy()
        y = y.copy()
        x[interior] += jx
        y[interior] += jy
    coords = np.stack([x, y], axis=1).astype(np.float64)

    sj, si = np.divmod(np.arange(n * n, dtype=np.int64), n)
    v00 = sj * n1 + si
    v10 = v00 + 1
    v01 = v00 + n1
    v11 = v01 + 1
    cells = np.empty((2 * n * n, 3), dtype=np.int64)
    cells[0::2] = np.stack([v00, v10, v11], axis=1)
    cells[1::2] = np.stack([v00, v11, v01], axis=1)

    u = rng.uniform(-1.0, 1.0, size=n1 * n1) if (rng is not None and draw_u) else None
    perm = None
    if permute:
        if rng is None:
            raise ValueError("permute needs an rng")
        perm = rng.permutation(n1 * n1)
        newc = np.empty_like(coords)
        newc[perm] = coords
        coords = newc
        if u is not None:
            newu = np.empty_like(u)
            newu[perm] = u
            u = newu
        cells = perm[cells]
    return TriMesh(n=n, coords=np.ascontiguousarray(coords),
                   cell_vertex=np.ascontiguousarray(cells.astype(np.int32)),
                   u=u, perm=perm)


def unit_cube_tet_p2(n, *, rng=None, draw_u=True):
    """``n^3`` unit cube split into 6 n^3 Kuhn tetrahedra with P2 numbering.

    Vertex (i, j, k) has id ``(k*(n+1) + j)*(n+1) + i``.  Hex (i, j, k) yields
    one tetrahedron per axis permutation p: v0 = corner, v1 = v0 + e_p0,
    v2 = v1 + e_p1, v3 = opposite corner.  P2 node ids are the vertex ids
    followed by ``nverts + edge`` where edges are numbered by ascending
    ``(min vertex, max vertex)``.  Local P2 nodes 4..9 follow ``TET_EDGES``.
    """
    n = int(n)
    if n < 1:
        raise ValueError("n must be >= 1")
    n1 = n + 1
    nv = n1 ** 3
    h = 1.0 / n
    kk, rem = np.divmod(np.arange(nv, dtype=np.int64), n1 * n1)
    jj, ii = np.divmod(rem, n1)
    coords = np.stack([ii * h, jj * h, kk * h], axis=1).astype(np.float64)

    hk, hrem = np.divmod(np.arange(n ** 3, dtype=np.int64), n * n)
    hj, hi = np.divmod(hrem, n)
    base = (hk * n1 + hj) * n1 + hi
    step = (1, n1, n1 * n1)
    perms = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
    cells = np.empty((n ** 3, 6, 4), dtype=np.int64)
    for t, p in enumerate(perms):
        v0 = base
        v1 = v0 + step[p[0]]
        v2 = v1 + step[p[1]]
        v3 = v2 + step[p[2]]
        cells[:, t, 0] = v0
        cells[:, t, 1] = v1
        cells[:, t, 2] = v2
        cells[:, t, 3] = v3
    cells = cells.reshape(-1, 4)

    ea = np.array([e[0] for e in TET_EDGES])
    eb = np.array([e[1] for e in TET_EDGES])
    a = cells[:, ea]
    b = cells[:, eb]
    lo = np.minimum(a, b)
    hi_ = np.maximum(a, b)
    keys = (lo * nv + hi_).ravel()
    uniq, inv = np.unique(keys, return_inverse=True)
    nedges = uniq.shape[0]
    edge_ids = inv.reshape(-1, 6)
    cell_node = np.concatenate([cells, nv + edge_ids], axis=1)

    u = None
    if rng is not None and draw_u:
        u = rng.uniform(-1.0, 1.0, size=(nv + nedges, 3))
    return TetP2Mesh(n=n, coords=coords,
                     cell_vertex=np.ascontiguousarray(cells.astype(np.int32)),
                     cell_node=np.ascontiguousarray(cell_node.astype(np.int32)),
                     nedges=int(nedges), u=u)


def batched_tri_meshes(count=512, nmin=16, nmax=128, *, rng=None) -> List[TriMesh]:
    """``count`` independent jittered unit squares with n ~ U{nmin..nmax}.

    The sizes are drawn first, then each mesh draws jitter x, jitter y and u.
    """
    if rng is None:
        rng = workload_rng(5)
    sizes = rng.integers(nmin, nmax + 1, size=count)
    return [unit_square_tri(int(s), rng=rng) for s in sizes]
