"""Tree-structured Parzen Estimator (Bergstra, Bardenet, Bengio and Kegl,
"Algorithms for Hyper-Parameter Optimization", NeurIPS 2011) over integer
parameters: the search the reference runs (examples/bayesian_search.py:63-187
creates an Optuna study with its default sampler, which is this algorithm).
Optuna is not installed here, so the sampler is written from the published
algorithm with the default settings that sampler documents:

- the first ``n_startup`` trials are uniform random;
- after that the trials are split by score: the best gamma(n) = min(ceil(0.1
  n), 25) form the "good" set, the rest the "bad" set;
- per parameter (independent, univariate), each set is a Parzen mixture of
  Gaussians truncated to [low - 0.5, high + 0.5] (an integer's interval):
  one component per observation plus a prior component (mean the interval's
  middle, width the interval, weight 1); each observation's width is the
  larger gap to its sorted neighbours (prior included), clipped to
  [range / min(100, 1 + n), range]; observation weights are 1 for the latest
  25 trials and a linear ramp from 1/n up to 1 for older ones;
- a proposal draws ``n_ei_candidates`` (24) values from the good mixture and
  keeps the one with the largest log l(x) - log g(x) (the expected
  improvement criterion), rounded to the nearest integer.

Batches: ``suggest(k)`` draws k proposals from the same model with
independent candidate draws, the way concurrent workers of one study sample
between updates; the simulator then judges all k at once.
"""
from __future__ import annotations

import math

import numpy as np

def _norm_cdf(x):
    from scipy.special import ndtr
    return ndtr(x)


def default_gamma(n: int) -> int:
    return min(int(math.ceil(0.1 * n)), 25)


def default_weights(n: int) -> np.ndarray:
    """1 for the latest 25 observations, a linear ramp from 1/n to 1 for the
    older ones (oldest first)."""
    if n == 0:
        return np.zeros(0)
    if n < 25:
        return np.ones(n)
    ramp = np.linspace(1.0 / n, 1.0, num=n - 25)
    return np.concatenate([ramp, np.ones(25)])


class Parzen:
    """Truncated Gaussian mixture of one parameter over [lo, hi] (already
    widened by half a step on each side for integers)."""

    def __init__(self, obs: np.ndarray, obs_weights: np.ndarray, lo: float, hi: float):
        n = obs.size
        prior_mu, prior_sigma = 0.5 * (lo + hi), hi - lo
        mus = np.concatenate([obs.astype(np.float64), [prior_mu]])
        w = np.concatenate([obs_weights.astype(np.float64), [1.0]])
        order = np.argsort(mus, kind="stable")
        sm = mus[order]
        if sm.size > 1:
            left = np.diff(sm, prepend=sm[0])
            right = np.diff(sm, append=sm[-1])
            left[0] = right[0]       # the ends use their one neighbour
            right[-1] = left[-1]
            s_sorted = np.maximum(left, right)
        else:
            s_sorted = np.array([prior_sigma])
        sig = np.empty_like(s_sorted)
        sig[order] = s_sorted
        sig_max = hi - lo
        sig_min = (hi - lo) / min(100.0, 1.0 + n)
        sig = np.clip(sig, sig_min, sig_max)
        sig[-1] = prior_sigma
        self.mu, self.sigma, self.lo, self.hi = mus, sig, lo, hi
        self.w = w / w.sum()
        a = (lo - self.mu) / self.sigma
        b = (hi - self.mu) / self.sigma
        self.mass = np.maximum(_norm_cdf(b) - _norm_cdf(a), 1e-300)

    def sample(self, rng: np.random.Generator, k: int) -> np.ndarray:
        comp = rng.choice(self.mu.size, size=k, p=self.w)
        out = np.empty(k)
        todo = np.arange(k)
        while todo.size:  # rejection into [lo, hi]
            x = rng.normal(self.mu[comp[todo]], self.sigma[comp[todo]])
            ok = (x >= self.lo) & (x <= self.hi)
            out[todo[ok]] = x[ok]
            todo = todo[~ok]
        return out

    def log_pdf(self, x: np.ndarray) -> np.ndarray:
        """log of the mixture density at x. The prior component (width = the
        whole interval) keeps the density of every x in the interval above
        w_prior * exp(-1/2) / (sigma sqrt(2 pi)), so the sum is formed
        directly (in float32, in row blocks) without underflow."""
        inv = (1.0 / self.sigma).astype(np.float32)
        mu = self.mu.astype(np.float32)
        coef = (self.w / (self.sigma * math.sqrt(2 * math.pi) * self.mass)).astype(np.float32)
        out = np.empty(x.size, np.float64)
        xs = x.astype(np.float32)
        step = max(1, (1 << 22) // max(1, mu.size))
        for a in range(0, x.size, step):
            z = (xs[a:a + step, None] - mu[None, :]) * inv[None, :]
            out[a:a + step] = np.log(np.exp(np.float32(-0.5) * z * z) @ coef)
        return out


class TPE:
    """Maximises a score over integer vectors within ranges[:, 0..1]
    (inclusive)."""

    def __init__(self, ranges: np.ndarray, seed: int = 0, n_startup: int = 10, n_ei_candidates: int = 24):
        self.ranges = np.asarray(ranges, np.int64)
        self.rng = np.random.default_rng(seed)
        self.n_startup = n_startup
        self.n_ei = n_ei_candidates
        self.X = np.zeros((0, self.ranges.shape[0]), np.int64)
        self.y = np.zeros(0)

    def tell(self, X: np.ndarray, y: np.ndarray) -> None:
        self.X = np.concatenate([self.X, np.atleast_2d(np.asarray(X, np.int64))])
        self.y = np.concatenate([self.y, np.atleast_1d(np.asarray(y, np.float64))])

    def _random(self, k: int) -> np.ndarray:
        lo, hi = self.ranges[:, 0], self.ranges[:, 1]
        return self.rng.integers(lo, hi + 1, size=(k, lo.size))

    def suggest(self, k: int = 1) -> np.ndarray:
        n = self.y.size
        if n < self.n_startup:
            return self._random(k)
        n_good = default_gamma(n)
        order = np.argsort(-self.y, kind="stable")  # best first
        good, bad = np.sort(order[:n_good]), np.sort(order[n_good:])  # trial order inside each set
        wg, wb = default_weights(good.size), default_weights(bad.size)
        out = np.empty((k, self.ranges.shape[0]), np.int64)
        for d in range(self.ranges.shape[0]):
            lo, hi = self.ranges[d, 0] - 0.5, self.ranges[d, 1] + 0.5
            lpe = Parzen(self.X[good, d], wg, lo, hi)
            gpe = Parzen(self.X[bad, d], wb, lo, hi)
            cand = lpe.sample(self.rng, k * self.n_ei).reshape(k, self.n_ei)
            ei = (lpe.log_pdf(cand.reshape(-1)) - gpe.log_pdf(cand.reshape(-1))).reshape(k, self.n_ei)
            best = cand[np.arange(k), np.argmax(ei, axis=1)]
            out[:, d] = np.clip(np.rint(best), self.ranges[d, 0], self.ranges[d, 1]).astype(np.int64)
        return out
