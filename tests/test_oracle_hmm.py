"""Pins the oracle's view-selection model: the two-state message recurrences
(computeMessageBackward/Forward, src/patchMatch.cu:669-715) against brute-force
enumeration of all 2^L hidden-state paths, and the CDF draw
(src/patchMatch.cu:917-961) on hand cases."""
import itertools

import numpy as np
import pytest

from oracle import pmx_oracle as O

# The reference's FP32 literals, at their FP32 values (src/patchMatch.cu:9-10, :673).
FIX = float(np.float32(0.999))
CHANGE = float(np.float32(1.0) - np.float32(0.999))
UNIFORM_X = float(np.float32(0.83))


def _brute_posteriors(costs, alpha):
    L = len(costs)
    e1 = np.exp(-0.5 * np.asarray(costs) ** 2 / alpha ** 2)       # view visible
    e0 = np.exp(-0.5 * UNIFORM_X ** 2 / alpha ** 2)              # occluded: uniform
    num = np.zeros(L)
    tot = 0.0
    for path in itertools.product((0, 1), repeat=L):
        p = 0.5 * (FIX + CHANGE)                                   # uniform prior, one transition
        for r, s in enumerate(path):
            if r > 0:
                p *= FIX if s == path[r - 1] else CHANGE
            p *= e1[r] if s else e0
        tot += p
        num += p * np.asarray(path)
    return num / tot


@pytest.mark.parametrize("L", [2, 3, 6, 10])
@pytest.mark.parametrize("alpha", [0.3, 0.45, 0.9])
def test_posterior_brute_force(L, alpha):
    costs = np.random.default_rng(L * 7 + int(alpha * 10)).uniform(0, 2, L)
    got = O.hmm_posteriors(costs, alpha)
    np.testing.assert_allclose(got[1:], _brute_posteriors(costs, alpha)[1:], rtol=1e-12, atol=1e-15)


def test_constant_column_fixed_point():
    # a long column of identical costs: the interior posterior is flat, and
    # a cost below the uniform level favours "visible", above it "occluded"
    for c, visible in [(0.1, True), (1.5, False)]:
        post = O.hmm_posteriors(np.full(40, c), 0.45)
        mid = post[15:25]
        assert np.ptp(mid) < 1e-12
        assert (mid[0] > 0.5) == visible


def test_pick_view_cases():
    cdf = np.array([0.2, 0.5, 1.0])
    assert O.pick_view(cdf, 1e-9) == 0
    assert O.pick_view(cdf, 0.2) == 0                       # u == cdf[i] takes i
    assert O.pick_view(cdf, np.nextafter(0.2, 1.0)) == 1
    assert O.pick_view(cdf, 0.5) == 1
    assert O.pick_view(cdf, 0.75) == 2
    assert O.pick_view(cdf, 1.0) == 2
    assert O.pick_view(np.array([0.0, 0.0, 1.0]), 2 ** -24) == 2   # zero-mass views are skipped
    assert O.pick_view(np.array([np.nan] * 4), 0.3) == 3
