"""Kernel recognition: the registry matches an element kernel by its
normalised source (tokens, without comments, layout or its function name),
so a reformatted or renamed copy of a hand-written family takes the HIP
entry point, and any change of meaning does not."""
import re

import numpy as np

from pyop2_amd import elements, op2
from pyop2_amd.registry import HandWrittenKernel, lookup

P1_NAME = "p1_laplace_residual"


def reformatted(src, old, new):
    """src with another function name, comments and a different layout."""
    out = src.replace(old, new)
    out = out.replace("{", "{ /* block */\n").replace(";", " ;  // step\n")
    return "// user copy\n" + re.sub(r"[ \t]+", "  ", out)


def test_digest_ignores_layout_comments_and_name():
    for family, params in (("p1_laplace_residual", {}), ("p2_elasticity_fused", {"lam": 1.0, "mu": 0.5})):
        src = elements.kernel_source(family, **params)
        copy = reformatted(src, family, "my_kernel")
        assert copy != src
        assert elements.code_digest(copy, "my_kernel") == elements.code_digest(src, family)


def test_digest_sees_changes_of_meaning():
    src = elements.kernel_source("p2_elasticity_residual", lam=1.0, mu=0.5)
    base = elements.code_digest(src, "p2_elasticity_residual")
    other = elements.kernel_source("p2_elasticity_residual", lam=2.0, mu=0.5)
    assert elements.code_digest(other, "p2_elasticity_residual") != base
    swapped = src.replace("+=", "-=", 1)
    assert swapped != src and elements.code_digest(swapped, "p2_elasticity_residual") != base
    # a character no C token covers is kept, not dropped
    odd = src + "\n@"
    assert elements.code_digest(odd, "p2_elasticity_residual") != base


def _p1_loop(code, name):
    n = 6
    cells = op2.Set(2 * n * n, "cells")
    verts = op2.Set((n + 1) ** 2, "verts")
    rng = np.random.default_rng(0)
    c2v = op2.Map(cells, verts, 3, rng.integers(0, (n + 1) ** 2, (2 * n * n, 3)), "c2v")
    X = op2.Dat(verts ** 2, rng.random(((n + 1) ** 2, 2)), name="X")
    u = op2.Dat(verts, rng.random((n + 1) ** 2), name="u")
    R = op2.Dat(verts, name="R")
    return op2.ParLoop(op2.Kernel(code, name), cells, R(op2.INC, c2v), X(op2.READ, c2v), u(op2.READ, c2v))


def test_user_written_copy_takes_the_hand_written_entry_point():
    """A reformatted, renamed copy of the P1 residual kernel, built with
    op2.Kernel directly, is recognised; a copy with one sign changed is not."""
    src = elements.kernel_source(P1_NAME)
    pl = _p1_loop(reformatted(src, P1_NAME, "laplace_res"), "laplace_res")
    ex = lookup(pl.global_kernel)
    assert isinstance(ex, HandWrittenKernel) and ex.family == P1_NAME
    wrong = reformatted(src, P1_NAME, "laplace_res").replace("+=", "-=", 1)
    assert lookup(_p1_loop(wrong, "laplace_res").global_kernel) is None


def _p1_typed_loop(dtypes):
    n = 6
    cells = op2.Set(2 * n * n, "cells")
    verts = op2.Set((n + 1) ** 2, "verts")
    rng = np.random.default_rng(0)
    c2v = op2.Map(cells, verts, 3, rng.integers(0, (n + 1) ** 2, (2 * n * n, 3)), "c2v")
    X = op2.Dat(verts ** 2, rng.random(((n + 1) ** 2, 2)).astype(dtypes[1]), dtypes[1], name="X")
    u = op2.Dat(verts, rng.random((n + 1) ** 2).astype(dtypes[2]), dtypes[2], name="u")
    R = op2.Dat(verts, dtype=dtypes[0], name="R")
    return op2.ParLoop(op2.Kernel(elements.kernel_source(P1_NAME), P1_NAME), cells, R(op2.INC, c2v),
                       X(op2.READ, c2v), u(op2.READ, c2v))


def test_registered_kernel_on_data_its_source_does_not_type_is_refused():
    """The registered sources take double arguments and the generated loop
    hands a kernel buffers of the data's type, so a registered kernel on
    float32 or mixed data that no hand-written entry point takes (the P1
    families have float64 entry points only) raises instead of running on
    mistyped buffers."""
    import pytest
    from pyop2_amd.exceptions import DataTypeError
    ok = lookup(_p1_typed_loop((np.float64,) * 3).global_kernel)
    assert isinstance(ok, HandWrittenKernel) and ok.variant == "f64"
    for dtypes in ((np.float32,) * 3, (np.float64, np.float32, np.float64)):
        with pytest.raises(DataTypeError, match="double arguments"):
            lookup(_p1_typed_loop(dtypes).global_kernel)
