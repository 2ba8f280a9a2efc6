"""The TPE sampler (axicor_amd/tpe.py) against the published algorithm's
rules and its purpose: it must find better points than uniform random
proposals with the same budget."""
import math

import numpy as np

from axicor_amd import tpe as T


def test_gamma_and_weights():
    assert [T.default_gamma(n) for n in (1, 9, 10, 11, 100, 250, 251, 10_000)] == [1, 1, 1, 2, 10, 25, 25, 25]
    assert T.default_weights(0).size == 0
    assert (T.default_weights(24) == 1).all()
    w = T.default_weights(30)
    assert w.size == 30 and (w[-25:] == 1).all()
    np.testing.assert_allclose(w[:5], np.linspace(1 / 30, 1, 5))


def test_parzen_widths_by_neighbour_gaps():
    """Observations 2, 5, 9 on [0, 10] with the prior at 5 (width 10): sorted
    means 2, 5, 5, 9; widths = larger neighbour gap, ends use their one gap,
    clipped to [10 / min(100, 4), 10]: 3 -> 3, 5's gaps 3 and 0 -> 3, prior 10,
    9 -> 4."""
    pe = T.Parzen(np.array([2, 5, 9]), np.ones(3), 0.0, 10.0)
    np.testing.assert_allclose(pe.mu, [2, 5, 9, 5])
    np.testing.assert_allclose(pe.sigma, [3, 3, 4, 10])
    np.testing.assert_allclose(pe.w, [0.25] * 4)
    # Clipping from below: one observation among many identical ones.
    pe = T.Parzen(np.full(50, 3.0), np.ones(50), 0.0, 10.0)
    assert np.isclose(pe.sigma[:-1].min(), 10.0 / 51)


def test_parzen_density_is_normalised_on_its_interval():
    rng = np.random.default_rng(1)
    pe = T.Parzen(rng.uniform(-0.5, 30.5, 40), T.default_weights(40), -0.5, 30.5)
    x = np.linspace(-0.5, 30.5, 20001)
    mass = np.trapezoid(np.exp(pe.log_pdf(x)), x)
    assert abs(mass - 1.0) < 1e-3
    s = pe.sample(rng, 5000)
    assert s.min() >= -0.5 and s.max() <= 30.5


def test_tpe_beats_random_search_on_a_quadratic():
    """12 integer parameters on the ranges of the search space, score
    -sum(((x - opt) / range)^2); 640 trials in batches of 64. Over five
    seeds TPE's best is better than random search's every time."""
    from axicor_amd.brain import PARAM_RANGES
    lo, hi = PARAM_RANGES[:, 0], PARAM_RANGES[:, 1]
    opt = lo + 0.3 * (hi - lo)

    def score(X):
        return -(((X - opt) / (hi - lo)) ** 2).sum(axis=1)

    for seed in range(5):
        t = T.TPE(PARAM_RANGES, seed=seed)
        for _ in range(10):
            X = t.suggest(64)
            t.tell(X, score(X))
        rng = np.random.default_rng(100 + seed)
        Xr = rng.integers(lo, hi + 1, size=(640, lo.size))
        assert t.y.max() > score(Xr).max(), seed
        assert np.all(t.X >= lo) and np.all(t.X <= hi)
