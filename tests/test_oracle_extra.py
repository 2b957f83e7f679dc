"""The oracle's Extra_LU+ and adaptive switch pinned by hand-derived answers (CPU only; the cases
and their derivations are in tests/extra_cases.py)."""
import numpy as np
import pytest

from oracle import binding

import extra_cases

EXTRA = 4   # HORA_ZOP_EXTRA_LU


@pytest.mark.parametrize("case", extra_cases.EXTRA_CASES, ids=lambda c: c[0])
def test_extra_lu_plus_known_answers(case):
    _, d, z, L, U, want = case
    out, _ = binding.zone_op(EXTRA, d, np.asarray([z], np.int32), L=np.asarray([L], np.int32),
                             U=np.asarray([U], np.int32))
    assert out[0].tolist() == want


def test_known_answers_reject_a_weaker_extrapolation():
    """The expected zones differ from the input wherever a clause fires, so an identity
    extrapolation, or one without a clause, cannot pass the test above."""
    changed = [c[0] for c in extra_cases.EXTRA_CASES if c[2] != c[5]]
    assert set(changed) == {"clause1", "clause2", "clause3", "clause4"}


def test_adaptive_switch_level():
    net, target, level = extra_cases.switch_model()
    r, st = binding.reach(net, target, 0)
    assert not r
    assert st["emode_switch_level"] == level
