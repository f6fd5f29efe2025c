"""Synthetic data of the benchmark configurations (seeded NumPy, host side).

Latent functions come from the GP prior: exactly (Cholesky of the prior
covariance) for small sets, and through random Fourier features (Rahimi &
Recht, 2007) -- cos features with frequencies drawn from the kernel's
spectral density: Gaussian for EQ, multivariate Cauchy for Matern12 -- when
an exact draw would need a large host Cholesky.  Ordinal labels are made as
in examples/classification.py:181-322: latent plus Gaussian noise, binned by
cutpoints.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy.special import ndtri

EQ = "EQ"
MATERN12 = "Matern12"


def _gram_np(kind, X, Z, ls, var, period=None):
    diff = X[:, None, :] - Z[None, :, :]
    if period is not None:
        q = 4.0 * np.sin(np.pi * diff / period) ** 2
    else:
        q = diff * diff
    r2 = np.sum(q / np.asarray(ls, dtype=np.float64) ** 2, axis=-1)
    return var * (np.exp(-0.5 * r2) if kind == EQ else np.exp(-np.sqrt(np.maximum(r2, 0.0))))


def prior_sample_exact(rng, kind, X, ls, var, period=None, jitter=1e-10):
    K = _gram_np(kind, X, X, ls, var, period) + jitter * np.eye(X.shape[0])
    L = np.linalg.cholesky(K)
    return L @ rng.standard_normal(X.shape[0])


def prior_sample_rff(rng, kind, X, ls, var, R=512):
    D = X.shape[1]
    ls = np.broadcast_to(np.asarray(ls, dtype=np.float64), (D,))
    g = rng.standard_normal((R, D))
    if kind == MATERN12:
        g = g / np.abs(rng.standard_normal((R, 1)))
    W = g / ls
    b = rng.uniform(0.0, 2.0 * np.pi, R)
    return math.sqrt(2.0 * var / R) * np.cos(X @ W.T + b).sum(axis=1)


def quantile_cutpoints(g, J):
    """Cutpoints between equal-count classes: midpoints of adjacent sorted
    responses at the class boundaries (examples/classification.py:249-272)."""
    gs = np.sort(g)
    n = gs.shape[0]
    cut = np.empty(J + 1)
    cut[0], cut[-1] = -np.inf, np.inf
    for j in range(1, J):
        k = (n * j) // J
        cut[j] = 0.5 * (gs[k - 1] + gs[k])
    return cut


def bin_labels(g, cut):
    return (np.searchsorted(cut[1:-1], g, side="right")).astype(np.float64)


@dataclass
class Problem:
    X: np.ndarray
    y: np.ndarray
    kind: str
    lengthscale: object
    variance: float
    lik: str  # "gaussian" or "probit"
    noise_std: float
    cutpoints: np.ndarray | None = None
    period: float | None = None
    Xtest: np.ndarray | None = None
    extra: dict = field(default_factory=dict)

    @property
    def N(self):
        return self.X.shape[0]

    @property
    def D(self):
        return self.X.shape[1]

    def likelihood_parameters(self):
        if self.lik == "gaussian":
            return (self.noise_std,)
        return (self.noise_std, self.cutpoints)


def regression_1d(seed=0, N=20, M=1000, noise_std=0.2):
    """Plain Gaussian regression, 1-D, kernel 1.0 * EQ().stretch(1).periodic(0.5):
    N training points X ~ U(-1, 1), M evenly spaced test points."""
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1.0, 1.0, (N, 1))
    Xt = np.linspace(-1.0, 1.0, M)[:, None]
    f = prior_sample_exact(rng, EQ, np.vstack([X, Xt]), 1.0, 1.0, period=0.5)
    y = f[:N] + noise_std * rng.standard_normal(N)
    return Problem(X, y, EQ, 1.0, 1.0, "gaussian", noise_std, period=0.5, Xtest=Xt,
                   extra={"f_test": f[N:]})


def ordinal(seed=1, N=8192, M=4096, D=2, J=3, kind=MATERN12, lengthscale=1.0, variance=1.0,
            noise_variance=0.4, x_dist="normal", exact=None):
    """Ordinal probit data: X ~ N(0, I) (or U(0,1)^D), latent from the prior,
    g = f + sqrt(noise_variance) * eps, J quantile-balanced classes."""
    rng = np.random.default_rng(seed)
    if x_dist == "normal":
        Xa = rng.standard_normal((N + M, D))
    else:
        Xa = rng.uniform(0.0, 1.0, (N + M, D))
    if exact is None:
        exact = N + M <= 3000
    if exact:
        f = prior_sample_exact(rng, kind, Xa, lengthscale, variance, jitter=1e-6)
    else:
        f = prior_sample_rff(rng, kind, Xa, lengthscale, variance)
    g = f + math.sqrt(noise_variance) * rng.standard_normal(N + M)
    cut = quantile_cutpoints(g[:N], J)
    y = bin_labels(g, cut)
    return Problem(Xa[:N], y[:N], kind, lengthscale, variance, "probit",
                   math.sqrt(noise_variance), cutpoints=cut, Xtest=Xa[N:],
                   extra={"y_test": y[N:], "g": g[:N]})


def ordinal_ard(seed=2, N=32768, M=4096, D=8, J=5, noise_variance=0.1):
    """Ordinal probit J=5, ARD EQ with lengthscales log-uniform in [0.3, 3],
    X ~ U(0,1)^D, cutpoints at standard-normal quantiles."""
    rng = np.random.default_rng(seed)
    ls = np.exp(rng.uniform(math.log(0.3), math.log(3.0), D))
    Xa = rng.uniform(0.0, 1.0, (N + M, D))
    f = prior_sample_rff(rng, EQ, Xa, ls, 1.0) if N + M > 3000 else \
        prior_sample_exact(rng, EQ, Xa, ls, 1.0, jitter=1e-8)
    g = f + math.sqrt(noise_variance) * rng.standard_normal(N + M)
    cut = np.concatenate([[-np.inf], ndtri(np.arange(1, J) / J), [np.inf]])
    y = bin_labels(g, cut)
    return Problem(Xa[:N], y[:N], EQ, ls, 1.0, "probit", math.sqrt(noise_variance),
                   cutpoints=cut, Xtest=Xa[N:], extra={"y_test": y[N:]})


def evidence_gradient(seed=3, N=8192, lik="gaussian", D=4, lengthscale=0.8, variance=1.5):
    """Evidence-gradient problem: X ~ N(0, I_D), latent from the EQ prior
    (lengthscale 0.8, signal variance 1.5); Gaussian targets with noise std
    0.2, or ordinal probit J=3 with noise variance 0.4 and quantile-balanced
    cutpoints."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((N, D))
    f = (prior_sample_exact(rng, EQ, X, lengthscale, variance, jitter=1e-8) if N <= 3000
         else prior_sample_rff(rng, EQ, X, lengthscale, variance))
    if lik == "gaussian":
        y = f + 0.2 * rng.standard_normal(N)
        return Problem(X, y, EQ, lengthscale, variance, "gaussian", 0.2)
    g = f + math.sqrt(0.4) * rng.standard_normal(N)
    cut = quantile_cutpoints(g, 3)
    return Problem(X, bin_labels(g, cut), EQ, lengthscale, variance, "probit", math.sqrt(0.4),
                   cutpoints=cut)


@dataclass
class Request:
    rid: int
    arrival: float
    problem: Problem


def serving_stream(seed=4, n=4096, rate=2000.0, M=256, Nmin=64, Nmax=4096):
    """Request stream: Poisson arrivals at `rate` per second; N log-uniform on
    [Nmin, Nmax], D uniform on {1..8}, X ~ U(-1,1)^D; Gaussian (noise std 0.2)
    or ordinal probit with J uniform on {2..5} (noise std 0.5, quantile
    cutpoints); EQ or Matern12 with lengthscale log-uniform on [0.2, 2];
    M test points per request."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for r in range(n):
        N = int(round(math.exp(rng.uniform(math.log(Nmin), math.log(Nmax)))))
        D = int(rng.integers(1, 9))
        kind = EQ if rng.random() < 0.5 else MATERN12
        ls = float(math.exp(rng.uniform(math.log(0.2), math.log(2.0))))
        X = rng.uniform(-1.0, 1.0, (N + M, D))
        f = prior_sample_rff(rng, kind, X, ls, 1.0, R=64)
        if rng.random() < 0.5:
            y = f[:N] + 0.2 * rng.standard_normal(N)
            p = Problem(X[:N], y, kind, ls, 1.0, "gaussian", 0.2, Xtest=X[N:])
        else:
            J = int(rng.integers(2, 6))
            g = f + 0.5 * rng.standard_normal(N + M)
            cut = quantile_cutpoints(g[:N], J)
            p = Problem(X[:N], bin_labels(g[:N], cut), kind, ls, 1.0, "probit", 0.5,
                        cutpoints=cut, Xtest=X[N:])
        out.append(Request(r, float(t[r]), p))
    return out
