"""Synthetic workload meshes: sizes of the workloads, orientation-free
positive volumes, Euler characteristic and the P2 node numbering rules."""
import numpy as np

from pyop2_amd.meshes import TET_EDGES, batched_tri_meshes, unit_cube_tet_p2, unit_square_tri, workload_rng


def test_triangle_sizes_and_areas():
    m = unit_square_tri(64, rng=workload_rng(1))
    assert m.ncells == 2 * 64 ** 2 and m.nverts == 65 ** 2
    x = m.coords[m.cell_vertex]
    det = (x[:, 1, 0] - x[:, 0, 0]) * (x[:, 2, 1] - x[:, 0, 1]) - (x[:, 2, 0] - x[:, 0, 0]) * (x[:, 1, 1] - x[:, 0, 1])
    assert (det > 0).all() and abs(det.sum() / 2 - 1.0) < 1e-12
    b = unit_square_tri(64)
    assert np.abs(m.coords - b.coords).max() <= 0.2 / 64 + 1e-15


def test_permutation_is_a_relabelling():
    a = unit_square_tri(16, rng=workload_rng(2))
    b = unit_square_tri(16, rng=workload_rng(2), permute=True)
    assert np.allclose(b.coords[b.perm], a.coords) and np.array_equal(b.cell_vertex, b.perm[a.cell_vertex])
    assert np.allclose(b.u[b.perm], a.u)


def test_tet_counts_and_volume():
    for n, (nedges, nnodes) in {32: (238688, 274625), 64: (1872064, 2146689)}.items():
        m = unit_cube_tet_p2(n)
        assert m.ncells == 6 * n ** 3 and m.nedges == nedges and m.nnodes == nnodes
        assert m.nverts - m.nedges + (m.ncells * 4 + 12 * n * n) // 2 - m.ncells == 1  # V-E+F-T=1
    m = unit_cube_tet_p2(4)
    x = m.coords[m.cell_vertex]
    J = np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], axis=2)
    vol = np.abs(np.linalg.det(J)) / 6
    assert abs(vol.sum() - 1.0) < 1e-12 and vol.min() > 0


def test_p2_numbering():
    m = unit_cube_tet_p2(3)
    cn, cv = m.cell_node, m.cell_vertex
    assert np.array_equal(cn[:, :4], cv)
    key = {}
    for e, (a, b) in enumerate(TET_EDGES):
        lo = np.minimum(cv[:, a], cv[:, b])
        hi = np.maximum(cv[:, a], cv[:, b])
        for k, node in zip(zip(lo, hi), cn[:, 4 + e]):
            assert key.setdefault(k, node) == node
    ordered = sorted(key)
    assert [key[k] for k in ordered] == list(range(m.nverts, m.nverts + m.nedges))


def test_batched_meshes_seeded():
    a = batched_tri_meshes(5)
    b = batched_tri_meshes(5)
    assert [m.n for m in a] == [m.n for m in b]
    assert all(16 <= m.n <= 128 for m in a)
    assert np.array_equal(a[3].coords, b[3].coords)
