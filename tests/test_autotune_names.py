"""Kernel-variant names and candidate lists of the autotuner (host logic,
no GPU): every candidate's name parses back to it, and the candidate list
follows the plan (second field set, material count)."""
from types import SimpleNamespace

import pytest

from rfx_amd import autotune as at


def _run(shape=(64, 40, 30), n_mat=10, two_sets=True, graded=False):
    prob = SimpleNamespace(shape=shape, graded=graded)
    return SimpleNamespace(prob=prob, flags=[0] * n_mat, sets=[0, 1] if two_sets else [0])


@pytest.mark.parametrize("two_sets", [False, True])
def test_names_round_trip(two_sets):
    for v in at.variants(_run(two_sets=two_sets)):
        name = at.variant_name(v)
        assert at.parse_variant(name) == tuple(v), name


def test_candidates_follow_the_plan():
    one = at.variants(_run(two_sets=False))
    assert {k for k, _ in one} == {0, 4, 5, 6}            # two-pass variants only
    two = at.variants(_run(two_sets=True))
    assert {1, 2, 3, 7, 8} <= {k for k, _ in two}
    # the wave-march and temporal-blocked steps stage at most 1024 material ids in LDS
    many = at.variants(_run(two_sets=True, n_mat=2000))
    assert not {7, 8} & {k for k, _ in many}
    # segment counts never exceed a quarter (an eighth for the two-step sweep) of the planes
    small = at.variants(_run(shape=(20, 8, 8), two_sets=True))
    assert all(seg <= 5 for k, seg in small if k in (2, 3, 7))
    assert all(seg <= 2 for k, seg in small if k == 8)


def test_known_names():
    assert at.variant_name((4, 1)) == "col2-flat"
    assert at.variant_name((0, 0)) == "two-pass-bands"
    assert at.variant_name((3, 32)) == "march14-32"
    assert at.variant_name((7, 0)) == "wmarch"
    assert at.parse_variant("wmarch-64") == (7, 64)
    assert at.parse_variant("col3-bands") == (6, 0)
    assert at.variant_name((8, 3)) == "tb2-3"
    assert at.parse_variant("tb2") == (8, 0)
