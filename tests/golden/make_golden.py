#!/usr/bin/env python3
"""Write the golden fixtures of the GP posterior path (small .npz files).

The reference cannot run here (its JAX / jaxopt / mlkernels dependencies are
absent), so the expected outputs come from the FP64 oracle, which
tests/test_oracle.py pins against independent checks (scikit-learn GP
regression, SciPy, torch autograd, finite differences, the reference's known
answers for h).  Each fixture stores the inputs and the oracle's outputs:

  regression_1d   Gaussian, kernel 1.0 * EQ().stretch(1).periodic(0.5), N=20,
                  1000 test points (the shape of examples/regression.py)
  ordinal_matern  ordinal probit J=3, Matern12, N=200, D=2
  ordinal_ard     ordinal probit J=5, ARD EQ, N=160, D=3
  vb_probit       VBGP, ordinal probit J=3, EQ, N=96, D=2
  likelihoods     log p and its first two derivatives on a grid of f for the
                  exact and the closed-form (BOUNDS) probit derivatives

Run from the repository root:  python tests/golden/make_golden.py
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import gp_oracle as O  # noqa: E402
from tests.helpers import oracle_spec, small_problem  # noqa: E402

TOL = 1e-5  # LaplaceGP / VBGP default tolerance (probit/approximators.py:62)


def laplace_case(X, y, lik_name, lp, spec, Xt):
    kind = O.GAUSSIAN if lik_name == "gaussian" else O.PROBIT
    K = O.gram(spec, X)
    w, it, _ = O.laplace_weight(K, y, kind, lp, tol=TOL)
    W, f = O.laplace_precision(K, y, kind, lp, w)
    mean, var = O.predict(spec, X, Xt, w, W)
    obj = O.objective_LA(spec, X, y, kind, lp, w)
    wc = O.laplace_weight_converged(K, y, kind, lp)
    g = O.laplace_objective_gradient(spec, X, y, kind, lp, wc)
    out = dict(weight=w, iterations=np.int64(it), precision=W, latent=f, mean=mean, var=var,
               objective=np.float64(obj), weight_converged=wc,
               objective_converged=np.float64(O.objective_LA(spec, X, y, kind, lp, wc)))
    for name, v in g.items():
        out["grad_" + name] = np.asarray(v, dtype=np.float64)
    return out


def spec_arrays(spec):
    return dict(kernel_kind=np.int64(spec.kind), variance=np.float64(spec.variance),
                lengthscale=np.asarray(spec.lengthscales, dtype=np.float64),
                period=(np.zeros(0) if spec.periods is None
                        else np.asarray(spec.periods, dtype=np.float64)))


def lik_arrays(lik_name, lp):
    d = dict(lik=np.array(lik_name), noise_std=np.float64(lp[0]))
    if lik_name != "gaussian":
        d["cutpoints"] = np.asarray(lp[1], dtype=np.float64)
    return d


def main():
    out_dir = HERE
    cases = {}
    # regression_1d: the shape of examples/regression.py (N=20, 1000 test points)
    rng = np.random.default_rng(0)
    X = rng.uniform(-1.0, 1.0, (20, 1))
    Xt = np.linspace(-1.0, 1.0, 1000)[:, None]
    spec = oracle_spec("EQ", 1.0, 1.0, 1, 0.5)
    Kall = O.gram(spec, np.vstack([X, Xt])) + 1e-10 * np.eye(1020)
    fall = np.linalg.cholesky(Kall) @ rng.standard_normal(1020)
    y = fall[:20] + 0.2 * rng.standard_normal(20)
    lp = (0.2,)
    cases["regression_1d"] = dict(X=X, y=y, Xt=Xt, **spec_arrays(spec), **lik_arrays("gaussian", lp),
                                  **laplace_case(X, y, "gaussian", lp, spec, Xt))
    # ordinal probit, Matern12, J = 3
    X, y, theta, lp, spec = small_problem(101, 200, 2, "probit", "Matern12", ls=1.0, var=1.0,
                                          noise=0.6, J=3)
    Xt = np.random.default_rng(102).normal(size=(128, 2))
    cases["ordinal_matern"] = dict(X=X, y=y, Xt=Xt, **spec_arrays(spec),
                                   **lik_arrays("probit", lp),
                                   **laplace_case(X, y, "probit", lp, spec, Xt))
    # ordinal probit, ARD EQ, J = 5
    rng = np.random.default_rng(103)
    X = rng.uniform(0.0, 1.0, (160, 3))
    ls = np.array([0.6, 1.1, 2.0])
    spec = oracle_spec("EQ", 1.0, ls, 3)
    f = np.linalg.cholesky(O.gram(spec, X) + 1e-8 * np.eye(160)) @ rng.standard_normal(160)
    g = f + np.sqrt(0.1) * rng.standard_normal(160)
    from scipy.special import ndtri
    cut = np.concatenate([[-np.inf], ndtri(np.arange(1, 5) / 5.0), [np.inf]])
    y = np.searchsorted(cut[1:-1], g, side="right").astype(np.float64)
    lp = (np.sqrt(0.1), cut)
    Xt = rng.uniform(0.0, 1.0, (64, 3))
    cases["ordinal_ard"] = dict(X=X, y=y, Xt=Xt, **spec_arrays(spec), **lik_arrays("probit", lp),
                                **laplace_case(X, y, "probit", lp, spec, Xt))
    # VBGP, probit J = 3
    X, y, theta, lp, spec = small_problem(104, 96, 2, "probit", "EQ", ls=0.9, var=1.2, noise=0.8)
    K = O.gram(spec, X)
    w, it, _ = O.vb_weight(K, y, O.PROBIT, lp, tol=TOL)
    cases["vb_probit"] = dict(X=X, y=y, **spec_arrays(spec), **lik_arrays("probit", lp),
                              weight=w, iterations=np.int64(it),
                              precision=O.vb_precision(lp, X.shape[0]),
                              objective=np.float64(O.objective_VB(spec, X, y, O.PROBIT, lp, w)))
    # likelihood derivative table
    fgrid = np.concatenate([np.linspace(-12.0, 12.0, 97), [-40.0, 40.0]])
    cut = np.array([-np.inf, -0.7, 0.15, 1.1, np.inf])
    lp = (0.65, cut)
    F, Y = np.meshgrid(fgrid, np.arange(4.0), indexing="ij")
    F, Y = F.ravel(), Y.ravel()
    tab = dict(f=F, y=Y, noise_std=np.float64(0.65), cutpoints=cut)
    for name, kind in (("exact", O.PROBIT), ("closed", O.PROBIT_CLOSED)):
        d = O.lik_derivs(kind, F, Y, lp)
        for key in ("ll", "d1", "d2"):
            tab[f"{name}_{key}"] = np.asarray(d[key], dtype=np.float64)
    yg = np.linspace(-2.0, 2.0, F.shape[0])
    d = O.lik_derivs(O.GAUSSIAN, F, yg, (0.3,))
    tab.update(gauss_y=yg, gauss_noise_std=np.float64(0.3),
               **{f"gauss_{k}": np.asarray(d[k]) for k in ("ll", "d1", "d2")})
    cases["likelihoods"] = tab
    for name, arrs in cases.items():
        path = os.path.join(out_dir, f"{name}.npz")
        np.savez_compressed(path, **arrs)
        print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
