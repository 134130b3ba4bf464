"""The segment-length tuners' decision rule (CPU): the fastest candidate wins,
but the library's own pick (0) stays unless beaten by more than the margin."""
from __future__ import annotations

from tenjin_amd import autotune


def test_choose_keeps_library_pick_within_margin():
    # 2D shock/inclusion FP64, a trial around mid-run: 8 rows 0.1 % slower
    assert autotune.choose({0: 0.00831, 8: 0.00841, 12: 0.00885}) == 0
    # a candidate 0.5 % faster is inside the 3 % margin
    assert autotune.choose({0: 1.000, 8: 0.995, 12: 1.02}) == 0


def test_choose_takes_a_clearly_faster_candidate():
    # axisymmetric explosion FP64: 8 and 12 rows about 7 % faster than the pick
    assert autotune.choose({0: 0.01027, 8: 0.00935, 12: 0.00936}) == 8
    assert autotune.choose({0: 1.0, 8: 0.99, 12: 0.90}) == 12
    # the margin is a parameter
    assert autotune.choose({0: 1.0, 8: 0.99}, margin=0.005) == 8


def test_choose_without_the_library_pick():
    assert autotune.choose({4: 2.0, 8: 1.5, 16: 1.7}) == 8


def test_level_candidates_include_the_library_pick():
    assert 0 in autotune.LEVEL_CANDIDATES and all(c >= 0 for c in autotune.LEVEL_CANDIDATES)
