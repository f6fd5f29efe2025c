"""Gradient of the VB objective (VBGP.value_and_grad through pgp_vb_grad)
against central finite differences of the oracle's objective_VB at tightly
converged fixed points: the value jax.value_and_grad gives for VBGP through
fixed_point_layer (probit/approximators.py:132-134,316-330,
probit/implicit/solvers.py:52-64, probit/implicit/VB.py:4-40)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import gp_oracle as O
from tests.helpers import lik_kind, make_prior, oracle_spec, small_problem

pytestmark = pytest.mark.gpu


def _oracle_objective(X, y, kind, lik, period, theta, lp):
    spec = oracle_spec(kind, theta[1], theta[0], X.shape[1], period)
    K = O.gram(spec, X)
    w = O.vb_weight(K, y, lik_kind(lik), lp, tol=1e-13, maxiter=100000)[0]
    return O.objective_VB(spec, X, y, lik_kind(lik), lp, w)


@pytest.mark.parametrize("lik,kind,period,noise", [("gaussian", "EQ", None, 0.6),
                                                   ("probit", "Matern12", None, 0.5),
                                                   ("probit", "EQ", 0.9, 0.7)])
def test_vb_gradient_matches_finite_differences(lik, kind, period, noise):
    import probit_amd as P
    X, y, theta, lp, _ = small_problem(40, 90, 2, lik, kind, noise=noise, period=period)
    ll = P.log_gaussian_likelihood if lik == "gaussian" else P.log_probit_likelihood
    gp = P.VBGP(data=(X, y), prior=make_prior(kind, period), log_likelihood=ll,
                tolerance=1e-12, maxiter=100000)
    print(f"vb grad {lik} {kind} {period}", flush=True)
    val, grads = gp.value_and_grad()((theta, lp))
    F = lambda th, lpp: _oracle_objective(X, y, kind, lik, period, th, lpp)  # noqa: E731
    ref = F(theta, lp)
    assert abs(val - ref) <= 1e-9 * abs(ref)
    (g_ls, g_var), g_lik = grads
    ls, var = theta
    h = 1e-5
    fd_ls = (F((ls * (1 + h), var), lp) - F((ls * (1 - h), var), lp)) / (2 * h * ls)
    fd_var = (F((ls, var * (1 + h)), lp) - F((ls, var * (1 - h)), lp)) / (2 * h * var)
    s = lp[0]
    lp_p = (s * (1 + h),) + tuple(lp[1:])
    lp_m = (s * (1 - h),) + tuple(lp[1:])
    fd_s = (F(theta, lp_p) - F(theta, lp_m)) / (2 * h * s)
    pairs = [(g_ls, fd_ls), (g_var, fd_var), (g_lik[0], fd_s)]
    if lik == "probit":
        cut = np.asarray(lp[1], dtype=np.float64)
        for k in range(1, len(cut) - 1):
            cp, cm = cut.copy(), cut.copy()
            cp[k] += h
            cm[k] -= h
            fd = (F(theta, (s, cp)) - F(theta, (s, cm))) / (2 * h)
            pairs.append((float(np.asarray(g_lik[1])[k]), fd))
        assert float(np.asarray(g_lik[1])[0]) == 0.0 and float(np.asarray(g_lik[1])[-1]) == 0.0
    for got, fd in pairs:
        assert abs(got - fd) <= 1e-6 * max(1.0, abs(fd)), (got, fd)


def test_vb_gradient_keeps_the_predict_state():
    """predict after value_and_grad equals predict right after the fit."""
    import probit_amd as P
    # sigma >= 0.5: the VB fixed-point iteration contracts
    X, y, theta, lp, _ = small_problem(41, 150, 2, "probit", "EQ", noise=0.6)
    gp = P.VBGP(data=(X, y), prior=make_prior("EQ"), log_likelihood=P.log_probit_likelihood,
                tolerance=1e-10, maxiter=2000)
    Xt = np.random.default_rng(42).normal(size=(70, 2))
    w, p = gp.approximate_posterior((theta, lp))
    assert gp.last_result.status == 0, gp.last_result.status  # converged
    m0, v0 = gp.predict(Xt, (theta, lp), w, p)
    m0, v0 = m0.cpu().numpy().copy(), v0.cpu().numpy().copy()
    gp.value_and_grad()((theta, lp))
    w, p = gp.approximate_posterior((theta, lp))
    gp._grad(*gp._structs((theta, lp)))
    m1, v1 = gp.predict(Xt, (theta, lp), w, p)
    assert np.array_equal(m1.cpu().numpy(), m0) and np.array_equal(v1.cpu().numpy(), v0)


@pytest.mark.parametrize("lik", ["gaussian", "probit"])
def test_vb_construct_is_the_oracle_map(lik):
    """VBGP.construct() is f_VB (probit/implicit/VB.py:4-16) at any weight, and
    the converged weight is its fixed point; predict afterwards is unchanged."""
    import probit_amd as P
    # sigma >= 0.5: the VB fixed-point iteration contracts (its iteration
    # matrix has eigenvalues up to |1 - 1/sigma| for a Gaussian likelihood)
    X, y, theta, lp, spec = small_problem(43, 130, 2, lik, "Matern12", noise=0.6)
    ll = P.log_gaussian_likelihood if lik == "gaussian" else P.log_probit_likelihood
    gp = P.VBGP(data=(X, y), prior=make_prior("Matern12"), log_likelihood=ll, tolerance=1e-11,
                maxiter=100000)
    f = gp.construct()
    w = np.random.default_rng(44).normal(size=130)
    got = f((theta, lp), w).cpu().numpy()
    ref = O.f_VB(O.gram(spec, X), y, lik_kind(lik), lp, w)
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))
    ws, ps = gp.approximate_posterior((theta, lp))
    Xt = np.random.default_rng(45).normal(size=(20, 2))
    m0, v0 = gp.predict(Xt, (theta, lp), ws, ps)
    m0, v0 = m0.cpu().numpy().copy(), v0.cpu().numpy().copy()
    fw = f((theta, lp), ws).cpu().numpy()
    assert np.max(np.abs(fw - ws.cpu().numpy())) <= 1e-9 * np.max(np.abs(fw))
    m1, v1 = gp.predict(Xt, (theta, lp), ws, ps)
    assert np.allclose(m1.cpu().numpy(), m0, rtol=1e-12, atol=0)
    assert np.allclose(v1.cpu().numpy(), v0, rtol=1e-12, atol=0)
