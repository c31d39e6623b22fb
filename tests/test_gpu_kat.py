"""The device par_loop (generated and hand-written paths) against the
reference's known answers: every ``kat_cases`` case through
``pyop2_amd.op2`` on the GPU, under each conflict-resolution strategy."""
import pytest

import kat_cases
from pyop2_amd import op2
from pyop2_amd.configuration import configuration

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("strategy", ["atomic", "coloured"])
@pytest.mark.parametrize("split", kat_cases.SPLITS, ids=["all", "nocore", "half"])
@pytest.mark.parametrize("case", kat_cases.LOOP_CASES, ids=lambda c: c.__name__)
def test_loop_cases(device, case, split, strategy):
    configuration["strategy"] = strategy
    case(op2, split)


@pytest.mark.parametrize("name,case", kat_cases.OTHER_CASES, ids=[n for n, _ in kat_cases.OTHER_CASES])
def test_other_cases(device, name, case):
    case(op2)
