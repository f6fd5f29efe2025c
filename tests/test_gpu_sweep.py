"""Batched objective + gradient over many hyper-parameter sets
(LaplaceGP.value_and_grad_batch, the lengthscale sweep
vmap(value_and_grad(nelbo)) of examples/classification.py:489-504) against
the one-set-at-a-time value_and_grad, and at one set against the oracle."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import gp_oracle as O
from tests.helpers import small_problem

pytestmark = pytest.mark.gpu


def _prior(kind):
    import probit_amd as P
    base = P.EQ if kind == "EQ" else P.Matern12

    def prior(theta):
        ls, var = theta
        return var * base().stretch(ls)

    return prior


@pytest.mark.parametrize("lik,kind,mode", [("probit", "EQ", "fp64"),
                                           ("gaussian", "Matern12", "fp64"),
                                           ("probit", "Matern12", "fp32")])
def test_sweep_matches_one_at_a_time(lik, kind, mode):
    import probit_amd as P
    X, y, theta, lp, spec = small_problem(21, 300, 2, lik, kind)
    ll = P.log_gaussian_likelihood if lik == "gaussian" else P.log_probit_likelihood
    gp = P.LaplaceGP(data=(X, y), prior=_prior(kind), log_likelihood=ll, tolerance=1e-10,
                     precision_mode=mode)
    plist = [((float(l), theta[1]), lp) for l in np.logspace(-0.7, 0.5, 9)]
    print(f"sweep {lik} {kind} {mode}: {len(plist)} sets", flush=True)
    vals, grads = gp.value_and_grad_batch()(plist)
    assert vals.shape == (len(plist),)
    tol = 1e-10 if mode == "fp64" else 1e-6
    for p, v, g in zip(plist, vals, grads):
        v1, g1 = gp.value_and_grad()(p)
        assert abs(v - v1) <= tol * abs(v1), (p[0], v, v1)
        (gl, gv), glik = g
        (gl1, gv1), glik1 = g1
        for a, b in ((gl, gl1), (gv, gv1), (glik[0], glik1[0])):
            assert abs(a - b) <= 1e3 * tol * max(1.0, abs(b)), (p[0], a, b)
    # one set against the oracle's objective at its converged weight
    (ls0, var0), lp0 = plist[4]
    spec0 = O.KernelSpec.make(O.EQ if kind == "EQ" else O.MATERN12, var0, ls0, 2)
    K = O.gram(spec0, X)
    kl = O.GAUSSIAN if lik == "gaussian" else O.PROBIT
    w = O.laplace_weight_converged(K, y, kl, lp0)
    ref = O.objective_LA(spec0, X, y, kl, lp0, w)
    assert abs(vals[4] - ref) <= (1e-8 if mode == "fp64" else 1e-5) * abs(ref)
