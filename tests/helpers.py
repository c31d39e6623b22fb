"""Shared helpers of the tests: op2 objects for the synthetic meshes and the
oracle's answer for the same inputs."""
import numpy as np

from oracle import seqwrap, sparsity as osp
from pyop2_amd import elements, op2


def use_strategy(name):
    """Select a par_loop strategy; "staged_atomic" is the staged strategy with
    device atomics for the nodes tiles share (no partial-sum pass)."""
    from pyop2_amd.configuration import configuration
    configuration["staged_partials"] = name != "staged_atomic"
    configuration["strategy"] = "staged" if name == "staged_atomic" else name


def relerr(a, b):
    """Normwise relative error ||a - b|| / ||b||."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


class TriProblem:
    """op2 objects of a P1 triangle mesh (optionally with a core/owned split)."""

    def __init__(self, mesh, core=None):
        self.m = mesh
        n = mesh.ncells
        sizes = n if core is None else (core, n, n)
        self.cells = op2.Set(sizes, "cells")
        self.verts = op2.Set(mesh.nverts, "verts")
        self.c2v = op2.Map(self.cells, self.verts, 3, mesh.cell_vertex, "c2v")
        self.X = op2.Dat(self.verts ** 2, mesh.coords, name="X")
        self.u = op2.Dat(self.verts, mesh.u, name="u")

    def sparsity(self):
        return op2.Sparsity((self.verts ** 1, self.verts ** 1), [(self.c2v, self.c2v, None)])


class TetProblem:
    """op2 objects of a P2 tetrahedral mesh."""

    def __init__(self, mesh, dtype=np.float64, core=None):
        self.m = mesh
        n = mesh.ncells
        sizes = n if core is None else (core, n, n)
        self.cells = op2.Set(sizes, "cells")
        self.verts = op2.Set(mesh.nverts, "verts")
        self.nodes = op2.Set(mesh.nnodes, "nodes")
        self.c2v = op2.Map(self.cells, self.verts, 4, mesh.cell_vertex, "c2v")
        self.c2n = op2.Map(self.cells, self.nodes, 10, mesh.cell_node, "c2n")
        self.X = op2.Dat(self.verts ** 3, mesh.coords.astype(dtype), dtype, name="X")
        self.u = op2.Dat(self.nodes ** 3, mesh.u.astype(dtype), dtype, name="u")

    def sparsity(self):
        return op2.Sparsity((self.nodes ** 3, self.nodes ** 3), [(self.c2n, self.c2n, None)])


def oracle_p1(family, mesh):
    """Oracle result of a P1 family on ``mesh`` (numpy)."""
    code = elements.kernel_source(family)
    cv = mesh.cell_vertex
    if family == "p1_lumped_mass":
        out = np.zeros(mesh.nverts)
        seqwrap.run(code, family, [seqwrap.DatArg(out, "INC", cv), seqwrap.DatArg(mesh.coords, "READ", cv)],
                    [(0, mesh.ncells)])
        return out
    if family == "p1_laplace_residual":
        out = np.zeros(mesh.nverts)
        seqwrap.run(code, family, [seqwrap.DatArg(out, "INC", cv), seqwrap.DatArg(mesh.coords, "READ", cv),
                                   seqwrap.DatArg(mesh.u, "READ", cv)], [(0, mesh.ncells)])
        return out
    if family == "p1_laplace_matrix":
        rp, ci = osp.pattern(mesh.nverts, mesh.nverts, [(cv, cv, mesh.ncells)])
        A = seqwrap.CSRMat(rp, ci, np.zeros(len(ci)))
        err = seqwrap.run(code, family, [seqwrap.MatArg(A, "INC", cv, cv), seqwrap.DatArg(mesh.coords, "READ", cv)],
                          [(0, mesh.ncells)])
        assert err == 0
        return A
    raise ValueError(family)


def oracle_p2(family, mesh, lam=1.0, mu=0.5, u=None, coords=None):
    code = elements.kernel_source(family, lam=lam, mu=mu)
    cv, cn = mesh.cell_vertex, mesh.cell_node
    X = mesh.coords if coords is None else coords
    uu = mesh.u if u is None else u
    if family == "p2_elasticity_residual":
        out = np.zeros((mesh.nnodes, 3))
        seqwrap.run(code, family, [seqwrap.DatArg(out, "INC", cn), seqwrap.DatArg(X, "READ", cv),
                                   seqwrap.DatArg(uu, "READ", cn)], [(0, mesh.ncells)])
        return out
    if family == "p2_elasticity_matrix":
        rp, ci = osp.pattern(mesh.nnodes, mesh.nnodes, [(cn, cn, mesh.ncells)])
        A = seqwrap.CSRMat(rp, ci, np.zeros((len(ci), 3, 3)), 3, 3, block=True)
        err = seqwrap.run(code, family, [seqwrap.MatArg(A, "INC", cn, cn, 3, 3), seqwrap.DatArg(X, "READ", cv)],
                          [(0, mesh.ncells)])
        assert err == 0
        return A
    raise ValueError(family)


def repeated_edge_nodes(mesh, fraction=0.05, seed=11):
    """``mesh`` with the last edge node of a ``fraction`` of its cells replaced
    by the cell's previous edge node, so those map rows name one node twice
    (the geometry, read through the vertex map, is unchanged).  A cell is
    only changed while the node it drops stays referenced by another cell,
    so every node keeps at least one incidence."""
    import dataclasses
    cn = mesh.cell_node.copy()
    refs = np.bincount(cn.ravel(), minlength=mesh.nnodes)
    rng = np.random.default_rng(seed)
    changed = 0
    for e in rng.permutation(mesh.ncells)[: int(fraction * mesh.ncells)]:
        if refs[cn[e, 9]] > 1:
            refs[cn[e, 9]] -= 1
            refs[cn[e, 8]] += 1
            cn[e, 9] = cn[e, 8]
            changed += 1
    assert changed > 0 and refs.min() > 0
    return dataclasses.replace(mesh, cell_node=cn)



def unreferenced_nodes(mesh, nholes=37, seed=5):
    """``mesh`` with ``nholes`` extra P2 nodes that no cell references,
    spread through the node numbering: their Mat rows hold only the
    diagonal block and no tile stages them."""
    import dataclasses
    rng = np.random.default_rng(seed)
    n = mesh.nnodes + nholes
    holes = np.sort(rng.choice(n, nholes, replace=False))
    keep = np.setdiff1d(np.arange(n), holes)        # new id of old node i
    u = np.empty((n, 3))
    u[keep] = mesh.u
    u[holes] = rng.standard_normal((nholes, 3))
    return dataclasses.replace(mesh, cell_node=keep[mesh.cell_node].astype(np.int32), u=u,
                               nedges=mesh.nedges + nholes), holes


def negative_entries(mesh, fraction=0.05, seed=13):
    """``mesh`` with the last edge node of a ``fraction`` of its cells
    replaced by -1 (no target: PETSc drops negative row and column ids), a
    cell only changed while the node it drops stays referenced."""
    import dataclasses
    cn = mesh.cell_node.copy()
    refs = np.bincount(cn.ravel(), minlength=mesh.nnodes)
    rng = np.random.default_rng(seed)
    for e in rng.permutation(mesh.ncells)[: int(fraction * mesh.ncells)]:
        if refs[cn[e, 9]] > 1:
            refs[cn[e, 9]] -= 1
            cn[e, 9] = -1
    assert (cn < 0).any() and refs.min() > 0
    return dataclasses.replace(mesh, cell_node=cn)
