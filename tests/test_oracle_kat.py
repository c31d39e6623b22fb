"""The oracle against the reference's known answers (CPU).

Runs every case of ``kat_cases`` through ``oracle_op2`` -- the oracle's
sequential wrapper (``oracle/seqwrap.py``) behind the op2 call forms -- over
the reference's core/owned split parametrisation."""
import pytest

import kat_cases
import oracle_op2


@pytest.mark.parametrize("split", kat_cases.SPLITS, ids=["all", "nocore", "half"])
@pytest.mark.parametrize("case", kat_cases.LOOP_CASES, ids=lambda c: c.__name__)
def test_loop_cases(case, split):
    case(oracle_op2, split)


@pytest.mark.parametrize("name,case", kat_cases.OTHER_CASES, ids=[n for n, _ in kat_cases.OTHER_CASES])
def test_other_cases(name, case):
    case(oracle_op2)


def test_planted_mistake_is_caught():
    """A wrapper that forgets to zero INC buffers must fail the INC case."""
    from oracle import seqwrap
    orig = seqwrap._generate

    def broken(code, name, args, ids, subset, headers=()):
        src = orig(code, name, args, ids, subset, headers)
        return src.replace("buf0[t_] = 0;", "buf0[t_] = 1;")
    seqwrap._generate = broken
    try:
        with pytest.raises(AssertionError):
            kat_cases.case_inc_unitset(oracle_op2, kat_cases.SPLITS[0])
    finally:
        seqwrap._generate = orig
