"""The binomial checkpointing plan of the adjoint (host code of libcfd_lbm_hip.so, no device call).

Pinned three ways: the closed form r*l - C(s+r, r-1) of the minimal forward-step count
(Griewank 1992), a brute-force dynamic programme over every split on small horizons (the true
optimum of reversing l steps with s free buffers), and the count obtained by walking the plan the
device code executes.
"""
import functools
from math import comb

import pytest

from cfd_solvers_amd import checkpoint_plan


@functools.lru_cache(maxsize=None)
def optimum(l, s):
    """Minimal forward steps to reverse l steps with the start state held and s free buffers."""
    if l <= 1:
        return 0
    if s == 0:
        return float("inf")
    return min(m + optimum(l - m, s - 1) + optimum(m, s) for m in range(1, l))


@pytest.mark.parametrize("s", [1, 2, 3, 4, 6])
def test_plan_is_optimal_on_small_horizons(s):
    for l in range(1, 90 if s < 4 else 60):
        fw, r, walked = checkpoint_plan(l, s)
        # + 1: the sweep also advances to the last state, which the objective reads.
        assert walked == fw == optimum(l, s) + 1, (l, s, fw, walked, optimum(l, s))
        assert comb(s + r, s) >= l and (r == 0 or comb(s + r - 1, s) < l)


def test_closed_form_at_scale():
    # 768^3 FP16 states of 17.2 GB: 12 free buffers for 10^4 steps -> every step recomputed at most 6 times.
    fw, r, walked = checkpoint_plan(10_000, 12)
    assert r == 6 and fw == walked == 6 * 10_000 - comb(18, 5) + 1
    # One buffer per step: no recomputation beyond the sweep.
    fw, r, walked = checkpoint_plan(2000, 1999)
    assert r == 1 and fw == walked == 2000
    # One buffer: quadratic.
    fw, r, walked = checkpoint_plan(50, 1)
    assert r == 49 and fw == walked == 50 * 49 // 2 + 1


def test_plan_rejects_empty_budgets():
    with pytest.raises(RuntimeError, match="snapshots"):
        checkpoint_plan(10, 0)
