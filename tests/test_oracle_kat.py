"""The oracle against the reference's known answers (CPU).

Runs every case of ``kat_cases`` through ``oracle_op2`` -- the oracle's
sequential wrapper (``oracle/seqwrap.py``) behind the op2 call forms -- over
the reference's core/owned split parametrisation."""
import re

import numpy as np
import pytest

import kat_cases
import oracle_op2


@pytest.mark.parametrize("split", kat_cases.SPLITS, ids=["all", "nocore", "half"])
@pytest.mark.parametrize("case", kat_cases.LOOP_CASES, ids=lambda c: c.__name__)
def test_loop_cases(case, split):
    case(oracle_op2, split)


@pytest.mark.parametrize("name,case", kat_cases.OTHER_CASES + kat_cases.MAP_CASES,
                         ids=[n for n, _ in kat_cases.OTHER_CASES + kat_cases.MAP_CASES])
def test_other_cases(name, case):
    case(oracle_op2)


@pytest.mark.parametrize("split", kat_cases.DIRECT_SPLITS, ids=["all", "nocore", "half", "halo"])
@pytest.mark.parametrize("case", kat_cases.DIRECT_CASES, ids=lambda c: c.__name__)
def test_direct_cases(case, split):
    case(oracle_op2, split)


@pytest.mark.parametrize("split", kat_cases.SPLITS, ids=["all", "nocore", "half"])
@pytest.mark.parametrize("case", kat_cases.REDUCTION_CASES, ids=lambda c: c.__name__)
def test_reduction_cases(case, split):
    case(oracle_op2, split)


@pytest.mark.parametrize("name,dtype,op,start,fill,want", kat_cases.TYPED_REDUCTIONS,
                         ids=[c[0] for c in kat_cases.TYPED_REDUCTIONS])
def test_typed_reductions(name, dtype, op, start, fill, want):
    kat_cases.case_reduce_typed(oracle_op2, kat_cases.SPLITS[2], dtype, op, start, fill, want)


def test_planted_composition_mistake_is_caught():
    """Dropping a middle map of the chain must fail the three-map case."""
    orig = oracle_op2.ComposedMap.table

    def wrong(self):
        tables = [m.table for m in self.maps_]
        out = []
        for n in range(self.iterset.total_size):
            e = int(tables[-1][n][0])
            out.append([int(v) for v in tables[0][e]])
        return np.array(out, dtype=np.int32).reshape(-1, self.arity)
    oracle_op2.ComposedMap.table = property(wrong)
    try:
        with pytest.raises(AssertionError):
            kat_cases.case_composed_three(oracle_op2, nested="none")
    finally:
        oracle_op2.ComposedMap.table = orig


def test_planted_permutation_mistake_is_caught():
    """Applying the inverse permutation must fail the permuted-map case."""
    orig = oracle_op2.PermutedMap.table

    def inverse(self):
        inv = np.argsort(self.permutation)
        return np.ascontiguousarray(self.map_.table[:, inv])
    oracle_op2.PermutedMap.table = property(inverse)
    try:
        with pytest.raises(AssertionError):
            kat_cases.case_permuted_map(oracle_op2)
    finally:
        oracle_op2.PermutedMap.table = orig


@pytest.mark.parametrize("split", kat_cases.SUBSET_SPLITS, ids=["all", "nocore", "half"])
@pytest.mark.parametrize("case", kat_cases.SUBSET_CASES, ids=lambda c: c.__name__)
def test_subset_cases(case, split):
    case(oracle_op2, split)


def test_subset_matrix():
    kat_cases.case_subset_matrix(oracle_op2)


@pytest.mark.parametrize("name,case", [c for c in kat_cases.MIXED_CASES if c[0] != "mixed_blockwise_sparsity"],
                         ids=[n for n, _ in kat_cases.MIXED_CASES if n != "mixed_blockwise_sparsity"])
def test_mixed_cases(name, case):
    case(oracle_op2)


@pytest.mark.parametrize("name,case", kat_cases.EXTRUDED_CASES, ids=[n for n, _ in kat_cases.EXTRUDED_CASES])
def test_extruded_cases(name, case):
    case(oracle_op2)


def test_planted_extrusion_mistake_is_caught():
    """Offsetting maps by the absolute layer instead of the layer above the
    column's bottom must fail the variable-layers case."""
    from oracle import seqwrap
    orig = seqwrap._generate

    def broken(*a, **k):
        return orig(*a, **k).replace("* (layer - bottom)", "* layer")
    seqwrap._generate = broken
    try:
        with pytest.raises(AssertionError):
            kat_cases.case_extruded_variable_layers(oracle_op2)
    finally:
        seqwrap._generate = orig


_REGION_PLANTS = [
    # the cell above named as the cell below
    ("above_is_below", lambda src: re.sub(r"\(layer - bottom \+ \(\w+\) / \d+\)", "(layer - bottom)", src),
     "case_extruded_interior_facets"),
    # the top region starting one layer low (two top cells)
    ("two_top_cells", lambda src: src.replace("cend - 1 > bottom ? cend - 1", "cend - 2 > bottom ? cend - 2"),
     "case_extruded_bottom_top"),
    # interior facets running one layer past the last facet
    ("facet_past_top", lambda src: src.replace("layer < cend - 1;", "layer < cend;"),
     "case_extruded_variable_regions"),
    # a periodic column missing the facet that joins its top cell to its bottom
    ("periodic_facet_missing", lambda src: src.replace("layer < cend; ++layer", "layer < cend - 1; ++layer")
     if "op2o_rem((layer - bottom + " in src else src, "case_periodic_facets"),
    # the offset quotient dropped (top vertex not wrapped)
    ("quotient_dropped", lambda src: re.sub(r"\(op2o_rem\(([^;]*?) \+ quo\d+\[[^\]]*\], ncell\) - "
                                           r"op2o_rem\(quo\d+\[[^\]]*\], ncell\)\)",
                                           r"op2o_rem(\1, ncell)", src),
     "case_periodic_vertices"),
    # no wrap at all on a periodic column
    ("no_wrap", lambda src: re.sub(r"op2o_rem\(([^,]*?), ncell\)", r"\1", src), "case_periodic_facets"),
]


@pytest.mark.parametrize("plant,case", [(p, c) for _, p, c in _REGION_PLANTS],
                         ids=[n for n, _, _ in _REGION_PLANTS])
def test_planted_region_mistake_is_caught(plant, case):
    """Each plausible iteration-region mistake, planted in the oracle's
    wrapper, fails one known-answer case."""
    from oracle import seqwrap
    orig = seqwrap._generate
    changed = []

    def broken(*a, **k):
        src = orig(*a, **k)
        out = plant(src)
        changed.append(out != src)
        return out
    seqwrap._generate = broken
    try:
        with pytest.raises(AssertionError):
            getattr(kat_cases, case)(oracle_op2)
    finally:
        seqwrap._generate = orig
    assert any(changed)


def test_planted_mistake_is_caught():
    """A wrapper that forgets to zero INC buffers must fail the INC case."""
    from oracle import seqwrap
    orig = seqwrap._generate

    def broken(*a, **k):
        return orig(*a, **k).replace("buf0[t_] = 0;", "buf0[t_] = 1;")
    seqwrap._generate = broken
    try:
        with pytest.raises(AssertionError):
            kat_cases.case_inc_unitset(oracle_op2, kat_cases.SPLITS[0])
    finally:
        seqwrap._generate = orig
