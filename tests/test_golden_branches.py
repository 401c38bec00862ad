"""Branch coverage of the golden fixtures: which branches of the farfield
boundary condition (FarfieldCellIds.cpp:55-113) and of AUSM's split Mach
numbers (AusmScheme.cpp:143-195, 777-833) the reference took in each run, as
tests/golden/make_golden.py counted them from the reference's own dumped
states (and checked each supersonic farfield ghost against its branch's
formula). Every branch must be taken by at least one fixture, so a wrong sign
or index in any of them fails a parity test."""
import pytest

from helpers import SMALL_FIXTURES, Golden

SUPERSONIC = [f for f in SMALL_FIXTURES if f.endswith("_sup")]


def branches(case):
    return Golden(case).meta["branches"]


def test_every_farfield_branch_is_taken():
    total = {}
    for case in SMALL_FIXTURES:
        for k, v in branches(case).items():
            total[k] = total.get(k, 0) + v
    for k in ("ff_sup_in", "ff_sup_out", "ff_sub_in", "ff_sub_out"):
        assert total[k] > 0, k


@pytest.mark.parametrize("case", SUPERSONIC)
def test_supersonic_fixture_takes_both_supersonic_farfield_branches(case):
    b = branches(case)
    assert b["ff_sup_in"] > 0 and b["ff_sup_out"] > 0, b


def test_ausm_split_mach_arms_are_taken():
    ausm = [f for f in SUPERSONIC if Golden(f).meta["flux"] == 1]
    assert len(ausm) >= 3 and any(Golden(f).meta["order"] == 1 for f in ausm)
    assert any(Golden(f).meta["order"] == 2 for f in ausm)
    assert any(Golden(f).meta["mesh"][3] == [2, 2, 2] for f in ausm)
    for f in ausm:
        b = branches(f)
        for arm in ("mach_L_le_m1", "mach_L_ge_1", "mach_R_le_m1", "mach_R_ge_1"):
            assert b[arm] > 0, (f, arm, b)
