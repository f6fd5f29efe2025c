"""Pins of the FP64 oracle against checks other than its own formulas.

* The reference's only known-answer test (probit/test/test_implicit.py:11-22).
* scikit-learn's independent GP regression (posterior mean/variance, log
  marginal likelihood and its gradient) for the Gaussian Laplace path.
* SciPy's normal CDF/PDF and torch-autograd derivatives for the
  likelihoods; brute-force loops for the Gram matrices; finite differences
  and closed-form special cases for Newton / objective / gradient / VB.
"""

import math

import numpy as np
import pytest
import scipy.linalg as sla
from scipy import stats

from oracle import gp_oracle as O
from tests.helpers import small_problem


# ---------------------------------------------------------------- h (a4) ---
def test_h_known_answers_from_reference_test():
    """probit/test/test_implicit.py:14-22."""
    with np.errstate(all="ignore"):
        assert np.isnan(O.h(0.0))
        assert O.h(np.inf) == 0.0
        assert O.h(-np.inf) == 0.0
        assert np.isclose(O.h(1.0), -65.0 / 6.0, rtol=1.0001)
        assert abs(O.h(1.0) - (-65.0 / 6.0)) < 1e-14  # exact value, stricter than the reference
        assert np.isnan(O.h_grad(0.0))
        assert O.h_grad(np.inf) == 0.0
        assert O.h_grad(-np.inf) == 0.0


def test_h_grad_is_derivative_of_h():
    x = np.array([0.7, 1.3, 2.5, -3.0, 8.0])
    fd = (O.h(x + 1e-6) - O.h(x - 1e-6)) / 2e-6
    assert np.allclose(O.h_grad(x), fd, rtol=1e-6)


def test_far_tail_series_approximates_normal_survival():
    """_Z_far_tails(z) ~ P(Z > z) to three decimals for z >= 4
    (probit/utilities.py:73-85 docstrings)."""
    z = np.array([4.0, 5.0, 6.5, 9.0])
    assert np.allclose(O._Z_far_tails(z), stats.norm.sf(z), rtol=1e-3)


# ---------------------------------------------------------------- Gram (a2) ---
def _gram_brute(kind, var, ls, per, X, Z):
    out = np.empty((X.shape[0], Z.shape[0]))
    for i in range(X.shape[0]):
        for j in range(Z.shape[0]):
            if per is None:
                r2 = sum(((X[i, d] - Z[j, d]) / ls[d]) ** 2 for d in range(X.shape[1]))
            else:
                # mlkernels feature map: (sin, cos)(2 pi x / p), then stretch
                r2 = 0.0
                for d in range(X.shape[1]):
                    a, b = 2 * math.pi * X[i, d] / per, 2 * math.pi * Z[j, d] / per
                    r2 += ((math.sin(a) - math.sin(b)) ** 2 + (math.cos(a) - math.cos(b)) ** 2) \
                        / ls[d] ** 2
            out[i, j] = var * (math.exp(-0.5 * r2) if kind == O.EQ else math.exp(-math.sqrt(r2)))
    return out


@pytest.mark.parametrize("kind", [O.EQ, O.MATERN12])
@pytest.mark.parametrize("per", [None, 0.6])
def test_gram_matches_brute_force(kind, per):
    rng = np.random.default_rng(0)
    X, Z = rng.normal(size=(7, 3)), rng.normal(size=(5, 3))
    ls = np.array([0.5, 1.0, 2.0])
    spec = O.KernelSpec(kind, 1.7, ls, None if per is None else np.full(3, per))
    assert np.allclose(O.gram(spec, X, Z), _gram_brute(kind, 1.7, ls, per, X, Z), rtol=1e-12)


def test_gram_invariants():
    rng = np.random.default_rng(1)
    X = rng.normal(size=(40, 2))
    spec = O.KernelSpec.make(O.EQ, 2.0, 0.9, 2)
    K = O.gram(spec, X)
    assert np.allclose(K, K.T)
    assert np.allclose(np.diag(K), 2.0)
    assert np.min(np.linalg.eigvalsh(K)) > -1e-10
    iso = O.gram(O.KernelSpec.make(O.EQ, 2.0, 0.9, 2), X)
    ard = O.gram(O.KernelSpec(O.EQ, 2.0, np.array([0.9, 0.9])), X)
    assert np.array_equal(iso, ard)


# ---------------------------------------------------------------- likelihoods (a3) ---
def test_gaussian_loglik_matches_scipy():
    rng = np.random.default_rng(2)
    f, y = rng.normal(size=50), rng.normal(size=50)
    assert np.allclose(O.log_gaussian_likelihood(f, y, (0.3,)),
                       stats.norm.logpdf(y, loc=f, scale=0.3), rtol=1e-13)


def test_probit_Z_matches_scipy_cdf():
    rng = np.random.default_rng(3)
    cut = np.array([-np.inf, -0.4, 0.3, np.inf])
    f = rng.normal(size=60)
    y = rng.integers(0, 3, size=60)
    ref = stats.norm.cdf((cut[y + 1] - f) / 0.7) - stats.norm.cdf((cut[y] - f) / 0.7)
    assert np.allclose(O.probit_Z(f, y, (0.7, cut)), ref, atol=1e-15)


def test_binary_probit_reduces_to_phi():
    """J = 2, cutpoints (-inf, 0, inf): P(y=1) = Phi(f / sigma) (R&W sec. 3.4)."""
    f = np.linspace(-3, 3, 13)
    lp = (0.8, np.array([-np.inf, 0.0, np.inf]))
    got = O.log_probit_likelihood(f, np.ones(13, dtype=int), lp)
    assert np.allclose(got, np.log(stats.norm.cdf(f / 0.8) + 1e-10), rtol=1e-13)


def _torch_probit_logp(f, y, s, cut):
    import torch
    lo, hi = cut[y], cut[y + 1]

    def ndtr(z):
        return 0.5 * (1 + torch.erf(z / math.sqrt(2)))

    c1 = torch.zeros_like(f) if np.isinf(lo) else ndtr((torch.tensor(lo) - f) / s)
    c2 = torch.ones_like(f) if np.isinf(hi) else ndtr((torch.tensor(hi) - f) / s)
    return torch.log(c2 - c1 + 1e-10)


@pytest.mark.parametrize("yv", [0, 1, 2])
def test_probit_derivatives_match_torch_autograd(yv):
    """Closed-form d1, d2, d3 and the sigma / cutpoint partials against torch
    float64 autograd of log(Phi(z2) - Phi(z1) + 1e-10)."""
    import torch
    cut = np.array([-np.inf, -0.4, 0.5, np.inf])
    for f0 in (-2.5, -0.3, 0.1, 1.7, 4.0):
        for s0 in (0.4, 1.1):
            ref = O.probit_derivs(np.array([f0]), np.array([yv]), (s0, cut), order=3)
            f = torch.tensor(f0, dtype=torch.float64, requires_grad=True)
            s = torch.tensor(s0, dtype=torch.float64, requires_grad=True)
            lp = _torch_probit_logp(f, yv, s, cut)
            d1, = torch.autograd.grad(lp, f, create_graph=True)
            d2, = torch.autograd.grad(d1, f, create_graph=True)
            d3, = torch.autograd.grad(d2, f, retain_graph=True)
            ls_, = torch.autograd.grad(lp, s, retain_graph=True)
            lfs, = torch.autograd.grad(d1, s, retain_graph=True)
            lffs, = torch.autograd.grad(d2, s, retain_graph=True)
            for got, want in ((ref["d1"], d1), (ref["d2"], d2), (ref["d3"], d3),
                              (ref["sigma"][0], ls_), (ref["sigma"][1], lfs),
                              (ref["sigma"][2], lffs)):
                assert np.allclose(got, float(want), rtol=1e-9, atol=1e-12)


def test_probit_cutpoint_partials_match_finite_differences():
    cut = np.array([-np.inf, -0.4, 0.5, np.inf])
    f = np.array([0.2, -1.0, 0.9])
    y = np.array([1, 1, 1])
    d = O.probit_derivs(f, y, (0.6, cut), order=3)
    h = 1e-6
    for j, key in ((1, "b_lo"), (2, "b_hi")):
        cp, cm = cut.copy(), cut.copy()
        cp[j] += h
        cm[j] -= h
        dp = O.probit_derivs(f, y, (0.6, cp))
        dm = O.probit_derivs(f, y, (0.6, cm))
        assert np.allclose(d[key][0], (dp["ll"] - dm["ll"]) / (2 * h), rtol=1e-6)
        assert np.allclose(d[key][1], (dp["d1"] - dm["d1"]) / (2 * h), rtol=1e-6)
        assert np.allclose(d[key][2], (dp["d2"] - dm["d2"]) / (2 * h), rtol=1e-5, atol=1e-8)


def test_closed_form_probit_agrees_inside_bounds():
    """Inside |z| < 1.3 the BOUNDS variant is the exact derivative (no floor)."""
    cut = np.array([-np.inf, -0.2, 0.3, np.inf])
    f = np.array([0.0, 0.05, 0.1])
    y = np.array([1, 1, 1])
    lp = (1.0, cut)
    ex = O.probit_derivs(f, y, lp)
    assert np.allclose(O.grad_log_probit_likelihood(f, y, lp), ex["d1"], rtol=1e-8)
    assert np.allclose(O.hessian_log_probit_likelihood(f, y, lp), ex["d2"], rtol=1e-8)
    # linear regime far in the tail: E = z1, V = -1/sigma^2
    f_far = np.array([-5.0])
    assert np.isclose(O.hessian_log_probit_likelihood(f_far, np.array([1]), lp)[0], -1.0)


# ---------------------------------------------------------------- Newton (a5-a8) ---
def test_gaussian_newton_is_exact_gp_solution():
    """Gaussian likelihood: the Laplace weight is (K + sigma^2 I)^-1 y (R&W Alg. 2.1)."""
    X, y, theta, lp, spec = small_problem(5, 60, 2, "gaussian")
    K = O.gram(spec, X)
    w, it, _ = O.laplace_weight(K, y, O.GAUSSIAN, lp, tol=1e-5)
    ref = sla.cho_solve(sla.cho_factor(K + lp[0] ** 2 * np.eye(60)), y)
    assert np.allclose(w, ref, rtol=1e-10)
    assert it == 2  # one exact step + one confirming step (jaxopt rule)


def test_bform_newton_equals_lu_newton():
    X, y, theta, lp, spec = small_problem(6, 80, 2, "probit")
    K = O.gram(spec, X)
    g1 = O.newton_map_lu(K, y, O.PROBIT, lp)
    g2 = O.newton_map_bform(K, y, O.PROBIT, lp)
    z = np.zeros(80)
    for _ in range(4):
        a, b = g1(z), g2(z)
        assert np.allclose(a, b, rtol=1e-10, atol=1e-12)
        z = a


def test_converged_weight_is_stationary():
    X, y, theta, lp, spec = small_problem(7, 90, 3, "probit", "Matern12")
    K = O.gram(spec, X)
    w = O.laplace_weight_converged(K, y, O.PROBIT, lp)
    d1 = O.lik_derivs(O.PROBIT, K @ w, y, lp)["d1"]
    assert np.max(np.abs(d1 - w)) < 1e-12


# ---------------------------------------------------------------- sklearn pins (a9-a11) ---
def _sklearn_gpr(kind, var, ls, noise, per=None, white=False):
    from sklearn.gaussian_process import GaussianProcessRegressor
    from sklearn.gaussian_process.kernels import (RBF, ConstantKernel, ExpSineSquared, Matern,
                                                  WhiteKernel)
    if per is not None:
        base = ExpSineSquared(length_scale=ls, periodicity=per)
    elif kind == O.EQ:
        base = RBF(length_scale=ls)
    else:
        base = Matern(length_scale=ls, nu=0.5)
    k = ConstantKernel(var) * base
    if white:
        k = k + WhiteKernel(noise ** 2)
        return GaussianProcessRegressor(kernel=k, alpha=0.0, optimizer=None)
    return GaussianProcessRegressor(kernel=k, alpha=noise ** 2, optimizer=None)


@pytest.mark.parametrize("kind,per", [(O.EQ, None), (O.MATERN12, None), (O.EQ, 0.8)])
def test_gaussian_predict_and_objective_match_sklearn(kind, per):
    rng = np.random.default_rng(8)
    X = rng.uniform(-1, 1, size=(50, 1))
    y = np.sin(4 * X[:, 0]) + 0.2 * rng.normal(size=50)
    Xt = np.linspace(-1, 1, 30)[:, None]
    spec = O.KernelSpec.make(kind, 1.3, 0.7, 1, per)
    lp = (0.25,)
    K = O.gram(spec, X)
    w, _, _ = O.laplace_weight(K, y, O.GAUSSIAN, lp, tol=1e-9)
    p, _ = O.laplace_precision(K, y, O.GAUSSIAN, lp, w)
    mean, var = O.predict(spec, X, Xt, w, p)
    gpr = _sklearn_gpr(kind, 1.3, 0.7, 0.25, per).fit(X, y)
    m_sk, s_sk = gpr.predict(Xt, return_std=True)
    assert np.allclose(mean, m_sk, rtol=1e-8, atol=1e-10)
    assert np.allclose(var, s_sk ** 2, rtol=1e-7, atol=1e-10)
    obj = O.objective_LA(spec, X, y, O.GAUSSIAN, lp, w)
    assert np.isclose(obj, -gpr.log_marginal_likelihood_value_, rtol=1e-10)
    cov = O.predict_covariance(spec, X, Xt, w, p)
    _, cov_sk = gpr.predict(Xt, return_cov=True)
    assert np.allclose(cov, cov_sk, rtol=1e-7, atol=1e-10)
    assert np.allclose(np.diag(cov), var, rtol=1e-10, atol=1e-13)


def test_gaussian_gradient_matches_sklearn():
    """d objective / d(variance, lengthscale, noise) vs sklearn's analytic
    gradient of the log marginal likelihood w.r.t. log-hyperparameters."""
    rng = np.random.default_rng(9)
    X = rng.normal(size=(45, 2))
    y = np.cos(X.sum(1)) + 0.2 * rng.normal(size=45)
    var, ls, noise = 1.4, 0.9, 0.3
    spec = O.KernelSpec.make(O.EQ, var, ls, 2)
    lp = (noise,)
    K = O.gram(spec, X)
    w = O.laplace_weight_converged(K, y, O.GAUSSIAN, lp)
    g = O.laplace_objective_gradient(spec, X, y, O.GAUSSIAN, lp, w)
    gpr = _sklearn_gpr(O.EQ, var, ls, noise, white=True)
    gpr.fit(X, y)
    lml, grad = gpr.log_marginal_likelihood(gpr.kernel_.theta, eval_gradient=True)
    # theta = log(var), log(ls), log(noise^2)
    assert np.isclose(-grad[0], g["variance"] * var, rtol=1e-8)
    assert np.isclose(-grad[1], (g["lengthscale0"] + g["lengthscale1"]) * ls, rtol=1e-8)
    assert np.isclose(-grad[2], g["noise_std"] * noise / 2.0, rtol=1e-8)


# ---------------------------------------------------------------- gradient (a11) ---
@pytest.mark.parametrize("lik", ["gaussian", "probit"])
def test_gradient_matches_finite_differences(lik):
    X, y, theta, lp, spec = small_problem(10, 50, 2, lik, "Matern12")

    def obj(var, ls, noise, cut=None):
        s = O.KernelSpec.make(O.MATERN12, var, ls, 2)
        l = (noise,) if lik == "gaussian" else (noise, cut)
        K = O.gram(s, X)
        return O.objective_LA(s, X, y, O.GAUSSIAN if lik == "gaussian" else O.PROBIT, l,
                              O.laplace_weight_converged(K, y, O.GAUSSIAN if lik == "gaussian"
                                                         else O.PROBIT, l))

    kind = O.GAUSSIAN if lik == "gaussian" else O.PROBIT
    K = O.gram(spec, X)
    g = O.laplace_objective_gradient(spec, X, y, kind, lp,
                                     O.laplace_weight_converged(K, y, kind, lp))
    h = 1e-5
    ls, var = theta
    cut = lp[1] if lik == "probit" else None
    fd_var = (obj(var + h, ls, lp[0], cut) - obj(var - h, ls, lp[0], cut)) / (2 * h)
    fd_ls = (obj(var, ls + h, lp[0], cut) - obj(var, ls - h, lp[0], cut)) / (2 * h)
    fd_n = (obj(var, ls, lp[0] + h, cut) - obj(var, ls, lp[0] - h, cut)) / (2 * h)
    assert np.isclose(g["variance"], fd_var, rtol=1e-6)
    assert np.isclose(g["lengthscale0"] + g["lengthscale1"], fd_ls, rtol=1e-6)
    assert np.isclose(g["noise_std"], fd_n, rtol=1e-6)


def test_gradient_matches_reference_implicit_differentiation():
    """The closed-form adjoint equals the reference's gradient structure:
    autodiff of objective_LA plus the fixed_point_layer VJP
    (probit/implicit/solvers.py:52-64), restated in torch float64."""
    from oracle.implicit_torch import implicit_gradient
    for lik in ("gaussian", "probit"):
        X, y, theta, lp, spec = small_problem(14, 40, 2, lik, "EQ")
        kind = O.GAUSSIAN if lik == "gaussian" else O.PROBIT
        K = O.gram(spec, X)
        w = O.laplace_weight_converged(K, y, kind, lp)
        g = O.laplace_objective_gradient(spec, X, y, kind, lp, w)
        gi = implicit_gradient(X, y, lik, theta[1], theta[0], lp, w)
        assert np.isclose(g["variance"], gi["variance"], rtol=1e-8)
        assert np.isclose(g["lengthscale0"] + g["lengthscale1"], gi["lengthscale"], rtol=1e-8)
        assert np.isclose(g["noise_std"], gi["noise_std"], rtol=1e-8)
        if lik == "probit":
            assert np.allclose(g["cutpoints"][1:-1], gi["cutpoints"][1:-1], rtol=1e-7)


# ---------------------------------------------------------------- VB (a12) ---
def test_vb_unit_noise_gaussian_is_exact_solution_and_nlml():
    """sigma = 1, Gaussian: the f_VB fixed point is (K + I)^-1 y and
    objective_VB equals the exact negative log marginal likelihood."""
    rng = np.random.default_rng(15)
    X = rng.normal(size=(40, 1))
    y = np.sin(2 * X[:, 0]) + rng.normal(size=40)
    spec = O.KernelSpec.make(O.EQ, 1.1, 0.8, 1)
    K = O.gram(spec, X)
    w, _, _ = O.vb_weight(K, y, O.GAUSSIAN, (1.0,), tol=1e-12)
    assert np.allclose(w, np.linalg.solve(K + np.eye(40), y), rtol=1e-9)
    obj = O.objective_VB(spec, X, y, O.GAUSSIAN, (1.0,), w)
    gpr = _sklearn_gpr(O.EQ, 1.1, 0.8, 1.0).fit(X, y)
    assert np.isclose(obj, -gpr.log_marginal_likelihood_value_, rtol=1e-9)


def test_predictive_distributions_sum_to_one():
    cut = np.array([-np.inf, -1.0, 0.0, 1.2, np.inf])
    mean, var = np.linspace(-2, 2, 9), np.linspace(0.1, 2.0, 9)
    P = O.probit_predictive_distributions((0.5, cut), mean, var)
    assert np.allclose(P.sum(axis=1), 1.0)
    s = np.sqrt(var + 0.25)
    assert np.allclose(P[:, 0], stats.norm.cdf((-1.0 - mean) / s))


def test_prior_sample_factor_and_covariance():
    """The oracle's prior sample is L z with L L^T = K + jitter I: the identity
    as z returns L itself (lower triangular, reconstructing K + jitter I), and
    the sample covariance of many draws is K + jitter I -- which L^T z (a
    transposed factor) or a dropped jitter would miss."""
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, size=(6, 2))
    spec = O.KernelSpec.make(O.EQ, 1.7, [0.6, 1.1], 2)
    jitter = 0.3
    K = O.gram(spec, X) + jitter * np.eye(6)
    L = O.prior_sample(spec, X, np.eye(6), jitter)
    assert np.allclose(np.triu(L, 1), 0.0)
    assert np.allclose(L @ L.T, K, rtol=1e-12, atol=1e-12)
    assert np.max(np.abs(L.T @ L - K)) > 0.1 * np.max(np.abs(K))  # the check below can tell
    z = rng.standard_normal((6, 200000))
    f = O.prior_sample(spec, X, z, jitter)
    C = f @ f.T / z.shape[1]
    assert np.max(np.abs(C - K)) < 0.03 * np.max(np.abs(K))


def test_prior_sample_of_uncorrelated_points_scales_z():
    """Points far apart under a short Matern12: K = variance I, so the sample
    is sqrt(variance + jitter) z exactly."""
    X = np.arange(5.0)[:, None] * 50.0
    spec = O.KernelSpec.make(O.MATERN12, 2.0, 0.01, 1)
    z = np.random.default_rng(6).standard_normal(5)
    f = O.prior_sample(spec, X, z, 1e-3)
    assert np.allclose(f, np.sqrt(2.0 + 1e-3) * z, rtol=1e-12, atol=0.0)


def test_calculate_metrics_hand_example():
    """examples/classification.py:325-341 on a 3-point, 2-class example worked
    by hand: argmax rows [1, 0, 1]; P[:, y] with y = [1, 1, 0] is the 3 x 3
    matrix of columns P[:, 1], P[:, 1], P[:, 0]."""
    P = np.array([[0.2, 0.8], [0.7, 0.3], [0.4, 0.6]])
    y = np.array([1, 1, 0])
    zo, mae, lpp, plik, y_pred = O.calculate_metrics(y, P)
    assert list(y_pred) == [1, 0, 1]
    assert zo == 2 / 3 and mae == 2 / 3
    expect = np.array([[0.8, 0.8, 0.2], [0.3, 0.3, 0.7], [0.6, 0.6, 0.4]])
    assert np.array_equal(plik, expect)
    assert np.isclose(lpp, 2 * np.log(0.8 * 0.3 * 0.6) + np.log(0.2 * 0.7 * 0.4), rtol=1e-14)
