"""Shared test helpers: problem builders that feed the same inputs to the HIP
path (through probit_amd) and to the oracle."""

from __future__ import annotations

import numpy as np

from oracle import gp_oracle as O

KINDS = {"EQ": O.EQ, "Matern12": O.MATERN12}


def oracle_spec(kind, variance, lengthscale, D, period=None):
    return O.KernelSpec.make(KINDS[kind], variance, lengthscale, D, period)


def lik_kind(name):
    return {"gaussian": O.GAUSSIAN, "probit": O.PROBIT}[name]


def make_prior(kind, period=None):
    """prior(theta) with theta = (lengthscale, variance), as in
    examples/regression.py:120-123."""
    from probit_amd import EQ, Matern12

    base = EQ if kind == "EQ" else Matern12

    def prior(theta):
        ls, var = theta
        k = var * base().stretch(ls)
        if period is not None:
            k = k.periodic(period)
        return k

    return prior


def small_problem(seed, N, D, lik, kind="EQ", ls=0.8, var=1.3, noise=0.3, J=3, period=None):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(N, D))
    f = np.sin(X.sum(axis=1)) + 0.3 * rng.normal(size=N)
    if lik == "gaussian":
        y = f
        lp = (noise,)
    else:
        g = f + noise * rng.normal(size=N)
        gs = np.sort(g)
        cut = np.concatenate([[-np.inf], [0.5 * (gs[(N * j) // J - 1] + gs[(N * j) // J])
                                          for j in range(1, J)], [np.inf]])
        y = np.searchsorted(cut[1:-1], g, side="right").astype(np.float64)
        lp = (noise, cut)
    theta = (ls, var)
    spec = oracle_spec(kind, var, ls, D, period)
    return X, y, theta, lp, spec


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b))))


def oracle_value_and_grad(X, y, kind, lik):
    """value_and_grad for the oracle at the FP64 Laplace fixed point of each
    parameter set ((lengthscale, variance), likelihood parameters), isotropic
    `kind` kernel: objective_LA and its closed-form total derivative, shaped
    like the parameters (the quantity jax.value_and_grad gives through
    fixed_point_layer, probit/approximators.py:132-134)."""
    lk = lik_kind(lik)

    def fn(params):
        (ls, var), lp = params
        spec = oracle_spec(kind, float(var), float(ls), X.shape[1])
        K = O.gram(spec, X)
        w = O.laplace_weight_converged(K, y, lk, lp)
        val = O.objective_LA(spec, X, y, lk, lp, w)
        g = O.laplace_objective_gradient(spec, X, y, lk, lp, w)
        g_ls = sum(g[f"lengthscale{d}"] for d in range(X.shape[1]))
        g_lik = (g["noise_std"],) if lik == "gaussian" else (g["noise_std"], g["cutpoints"])
        return float(val), ((g_ls, g["variance"]), g_lik)

    return fn
