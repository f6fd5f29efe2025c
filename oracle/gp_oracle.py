"""FP64 NumPy restatement of probit's GP posterior path (test infrastructure).

This module is the oracle: tests compare the HIP path against it.  It is never
imported by the product package.  Every function cites the reference
file:line it restates (paths relative to the reference repository root).

Conventions (shared with the product path):
  * X is (N, D) float64, y is (N,) float64 (targets, or class indices 0..J-1).
  * A kernel is a ``KernelSpec``: variance * kappa(r^2) with
        r^2 = sum_d q_d / lengthscale_d^2,
        q_d = (x_d - x'_d)^2                       (plain stretch)
        q_d = 4 sin^2(pi (x_d - x'_d) / period_d)  (periodic feature map),
    kappa = exp(-r^2/2) (EQ) or exp(-r) (Matern12).  This is mlkernels'
    ``s * EQ().stretch(l).periodic(p)`` composition (a third-party library the
    reference calls at probit/approximators.py:171-174; its semantics are
    restated from its published documentation and pinned by brute-force and
    closed-form checks in tests/test_oracle.py).
  * Likelihood parameters are ``(noise_std,)`` (Gaussian) or
    ``(noise_std, cutpoints)`` with cutpoints of length J+1 including the
    -inf / +inf ends (probit/utilities.py:47-53).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla
from scipy.special import erf

SQRT2 = math.sqrt(2.0)
OVER_SQRT_2PI = 1.0 / math.sqrt(2.0 * math.pi)  # probit/utilities.py:10
LOG_OVER_SQRT_2PI = math.log(OVER_SQRT_2PI)  # probit/utilities.py:11
PROBIT_FLOOR = 1e-10  # probit/utilities.py:57
BOUNDS = {"single": [1.3, 1.8, 2.3], "double": [2.3, 3.6, 4.8]}  # probit/utilities.py:15

EQ = 0
MATERN12 = 1


# ----------------------------------------------------------------------------
# Kernels (mlkernels semantics, reached via prior(theta)(X):
# probit/implicit/Laplace.py:7,21,24, probit/approximators.py:171-174)
# ----------------------------------------------------------------------------
@dataclass
class KernelSpec:
    kind: int  # EQ or MATERN12
    variance: float
    lengthscales: np.ndarray  # (D,)
    periods: np.ndarray | None = None  # (D,) or None for non-periodic

    @staticmethod
    def make(kind, variance, lengthscale, D, period=None):
        ls = np.broadcast_to(np.asarray(lengthscale, dtype=np.float64), (D,)).copy()
        per = None
        if period is not None:
            per = np.broadcast_to(np.asarray(period, dtype=np.float64), (D,)).copy()
        return KernelSpec(kind, float(variance), ls, per)


def _q(spec: KernelSpec, X: np.ndarray, Z: np.ndarray) -> np.ndarray:
    """Per-dimension squared feature distances q_d, shape (N, M, D)."""
    diff = X[:, None, :] - Z[None, :, :]
    if spec.periods is None:
        return diff * diff
    s = np.sin(np.pi * diff / spec.periods)
    return 4.0 * s * s


def gram(spec: KernelSpec, X: np.ndarray, Z: np.ndarray | None = None) -> np.ndarray:
    """Gram matrix K_ij = variance * kappa(r_ij^2).

    Restates the kernel evaluation ``B.dense(prior(theta)(X))``
    (probit/implicit/Laplace.py:7) and ``kernel(X, X_test)``
    (probit/approximators.py:173).
    """
    Z = X if Z is None else Z
    # r^2 accumulated one input dimension at a time (N x M memory, not N x M x D)
    r2 = np.zeros((X.shape[0], Z.shape[0]))
    for d in range(X.shape[1]):
        diff = X[:, d, None] - Z[None, :, d]
        if spec.periods is None:
            q = diff * diff
        else:
            sn = np.sin(np.pi * diff / spec.periods[d])
            q = 4.0 * sn * sn
        r2 += q / spec.lengthscales[d] ** 2
    if spec.kind == EQ:
        return spec.variance * np.exp(-0.5 * r2)
    return spec.variance * np.exp(-np.sqrt(np.maximum(r2, 0.0)))


def gram_elwise(spec: KernelSpec, Xt: np.ndarray) -> np.ndarray:
    """kernel.elwise(X_test, X_test): prior variance (probit/approximators.py:172)."""
    return np.full(Xt.shape[0], spec.variance)


def gram_param_derivs(spec: KernelSpec, X: np.ndarray):
    """dK/dtheta for theta = (variance, lengthscale_d..., period_d...).

    Returns a list of (name, N x N array).  Used by the closed-form adjoint.
    """
    q = _q(spec, X, X)
    ls = spec.lengthscales
    r2 = np.sum(q / ls**2, axis=-1)
    out = []
    if spec.kind == EQ:
        kap = np.exp(-0.5 * r2)
        dk_dr2 = -0.5 * spec.variance * kap
    else:
        r = np.sqrt(np.maximum(r2, 0.0))
        kap = np.exp(-r)
        with np.errstate(divide="ignore", invalid="ignore"):
            dk_dr2 = np.where(r > 0, -spec.variance * kap / (2.0 * r), 0.0)
    out.append(("variance", kap))
    for d in range(X.shape[1]):
        dr2 = -2.0 * q[:, :, d] / ls[d] ** 3
        out.append((f"lengthscale{d}", dk_dr2 * dr2))
    if spec.periods is not None:
        diff = X[:, None, :] - X[None, :, :]
        for d in range(X.shape[1]):
            P = spec.periods[d]
            dq = -4.0 * np.sin(2.0 * np.pi * diff[:, :, d] / P) * np.pi * diff[:, :, d] / P**2
            out.append((f"period{d}", dk_dr2 * dq / ls[d] ** 2))
    return out


# ----------------------------------------------------------------------------
# Likelihoods (probit/utilities.py)
# ----------------------------------------------------------------------------
def ndtr(z):
    """Normal CDF in the reference's erf form (probit/utilities.py:18-19)."""
    return 0.5 * (1.0 + erf(z / SQRT2))


def norm_z_pdf(z):
    """probit/utilities.py:22-23."""
    return OVER_SQRT_2PI * np.exp(-0.5 * z * z)


def log_gaussian_likelihood(f, y, lik):
    """log N(y; f, sigma^2) (probit/utilities.py:60-70)."""
    sigma = float(lik[0])
    z = (f - y) / sigma
    return LOG_OVER_SQRT_2PI - z * z / 2.0 - math.log(sigma)


def _probit_z(f, y, lik):
    """z-scores of the class interval, with the reference's +-inf guards
    (probit/utilities.py:47-53, 195-229)."""
    sigma = float(lik[0])
    cut = np.asarray(lik[1], dtype=np.float64)
    yi = np.asarray(y).astype(np.int64)
    b_lo = cut[yi]
    b_hi = cut[yi + 1]
    lo_inf = b_lo == -np.inf
    hi_inf = b_hi == np.inf
    z1 = np.where(lo_inf, 0.0, (np.where(lo_inf, 0.0, b_lo) - f) / sigma)
    z2 = np.where(hi_inf, 0.0, (np.where(hi_inf, 0.0, b_hi) - f) / sigma)
    return sigma, z1, z2, lo_inf, hi_inf


def probit_Z(f, y, lik):
    """Class probability Z = Phi(z2) - Phi(z1) (probit/utilities.py:195-229)."""
    _, z1, z2, lo_inf, hi_inf = _probit_z(f, y, lik)
    c1 = np.where(lo_inf, 0.0, ndtr(z1))
    c2 = np.where(hi_inf, 1.0, ndtr(z2))
    return c2 - c1


def log_probit_likelihood(f, y, lik):
    """log(Z + 1e-10) (probit/utilities.py:56-57)."""
    return np.log(probit_Z(f, y, lik) + PROBIT_FLOOR)


def gaussian_derivs(f, y, lik, order=2):
    """Exact derivatives of log_gaussian_likelihood w.r.t. f (the values JAX
    autodiff yields at probit/approximators.py:92-95).

    Returns dict with ll, d1, d2 and, for order 3, d3 plus the
    likelihood-parameter partials (key 'sigma': (L_p, L_fp, L_ffp)).
    """
    sigma = float(lik[0])
    r = y - f
    out = {
        "ll": log_gaussian_likelihood(f, y, lik),
        "d1": r / sigma**2,
        "d2": np.full_like(f, -1.0 / sigma**2),
    }
    if order >= 3:
        out["d3"] = np.zeros_like(f)
        out["sigma"] = (r * r / sigma**3 - 1.0 / sigma, -2.0 * r / sigma**3,
                        np.full_like(f, 2.0 / sigma**3))
    return out


def probit_derivs(f, y, lik, order=2):
    """Exact derivatives of log(Phi(z2) - Phi(z1) + 1e-10) w.r.t. f, sigma and
    the two cutpoints of each datum's class interval.

    These are the values JAX autodiff of probit/utilities.py:56-57 yields
    (probit/approximators.py:92-95 builds grad and grad-of-grad).  Infinite
    interval ends contribute zero (the jnp.where guards at
    probit/utilities.py:220-228 make their autodiff branches constant).
    """
    sigma, z1, z2, lo_inf, hi_inf = _probit_z(f, y, lik)
    c1 = np.where(lo_inf, 0.0, ndtr(z1))
    c2 = np.where(hi_inf, 1.0, ndtr(z2))
    Ze = (c2 - c1) + PROBIT_FLOOR
    p1 = np.where(lo_inf, 0.0, norm_z_pdf(z1))
    p2 = np.where(hi_inf, 0.0, norm_z_pdf(z2))
    Zf = (p1 - p2) / sigma
    Zff = (z1 * p1 - z2 * p2) / sigma**2
    L1 = Zf / Ze
    L2 = Zff / Ze - L1 * L1
    out = {"ll": np.log(Ze), "d1": L1, "d2": L2}
    if order >= 3:
        Zfff = ((z1 * z1 - 1.0) * p1 - (z2 * z2 - 1.0) * p2) / sigma**3
        out["d3"] = Zfff / Ze - 3.0 * L1 * L2 - L1**3

        def param(Zp, Zfp, Zffp):
            Lp = Zp / Ze
            Lfp = Zfp / Ze - L1 * Lp
            Lffp = Zffp / Ze - Zff * Zp / Ze**2 - 2.0 * L1 * Lfp
            return Lp, Lfp, Lffp

        out["sigma"] = param(
            (z1 * p1 - z2 * p2) / sigma,
            ((z1 * z1 - 1.0) * p1 - (z2 * z2 - 1.0) * p2) / sigma**2,
            ((z1**3 - 3.0 * z1) * p1 - (z2**3 - 3.0 * z2) * p2) / sigma**3,
        )
        out["b_lo"] = param(-p1 / sigma, -z1 * p1 / sigma**2, (1.0 - z1 * z1) * p1 / sigma**3)
        out["b_hi"] = param(p2 / sigma, z2 * p2 / sigma**2, -(1.0 - z2 * z2) * p2 / sigma**3)
    return out


# --- optional closed-form probit derivatives (probit/utilities.py:37-192) ---
def h(x):
    """Tail-series polynomial (probit/utilities.py:37-44)."""
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return -1.0 * x**-2 + 5.0 / 2.0 * x**-4 - 37.0 / 3.0 * x**-6


def h_grad(x):
    """d h / d x (pins test_implicit.py:20-22)."""
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return 2.0 * x**-3 - 10.0 * x**-5 + 74.0 * x**-7


def _Z_far_tails(z):
    """probit/utilities.py:83-85."""
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        return OVER_SQRT_2PI / z * np.exp(-0.5 * z**2 + h(z))


def _Z_tails(z1, z2):
    """probit/utilities.py:73-80."""
    return _Z_far_tails(z1) - _Z_far_tails(z2)


def norm_cdf(x):
    """probit/utilities.py:31-34."""
    x = np.asarray(x, dtype=np.float64)
    _x = np.where(np.isinf(x), 1.0, x)
    extrema = np.where(x == np.inf, 1.0, 0.0)
    return np.where(np.isinf(x), extrema, ndtr(_x))


def _safe_Z(f, y, lik, ub, ub2, ub3):
    """Piecewise class probability (probit/utilities.py:88-148)."""
    cut = np.asarray(lik[1], dtype=np.float64)
    yi = np.asarray(y).astype(np.int64)
    c_hi = cut[yi + 1]
    c_lo = cut[yi]
    sigma = float(lik[0])
    _b = np.where(c_hi == np.inf, 0.0, c_hi)
    _a = np.where(c_lo == -np.inf, 0.0, c_lo)
    z2s = np.where(c_hi == np.inf, np.inf, (_b - f) / sigma)
    z1s = np.where(c_lo == -np.inf, -np.inf, (_a - f) / sigma)
    SAFE = 1.0
    Z = norm_cdf(z2s) - norm_cdf(z1s)
    _z1s = np.where((ub < z1s) & (z1s <= ub2), z1s, SAFE)
    __z2s = np.where(ub < z1s, z2s, SAFE)
    _z2s = np.where((-ub2 <= z2s) & (z2s < -ub), z2s, SAFE)
    __z1s = np.where(-ub > z2s, z1s, SAFE)
    with np.errstate(all="ignore"):
        Z = np.where(z1s > ub, _Z_tails(_z1s, __z2s), Z)
        Z = np.where(z2s < -ub, _Z_tails(__z1s, _z2s), Z)
        _z1s = np.where((ub2 < np.abs(z1s)) & (np.abs(z1s) < ub3), z1s, SAFE)
        _z2s = np.where((ub2 < np.abs(z2s)) & (np.abs(z2s) < ub3), z2s, SAFE)
        Z = np.where(z1s > ub2, _Z_far_tails(_z1s), Z)
        Z = np.where(z2s < -ub2, _Z_far_tails(-_z2s), Z)
    Z = np.where(z1s >= ub3, SAFE, Z)
    Z = np.where(z2s <= -ub3, SAFE, Z)
    return Z, z1s, z2s


def _norm_pdf_inf(z):
    with np.errstate(over="ignore", invalid="ignore"):
        return np.where(np.isinf(z), 0.0, norm_z_pdf(np.where(np.isinf(z), 0.0, z)))


def grad_log_probit_likelihood(f, y, lik, single_precision=True):
    """probit/utilities.py:151-169."""
    ub, ub2, ub3 = BOUNDS["single" if single_precision else "double"]
    sigma = float(lik[0])
    Z, z1s, z2s = _safe_Z(f, y, lik, ub, ub2, ub3)
    with np.errstate(all="ignore"):
        E = (_norm_pdf_inf(z1s) - _norm_pdf_inf(z2s)) / Z
    E = np.where(z1s > ub3, z1s, E)
    E = np.where(z2s < -ub3, z2s, E)
    return E / sigma


def hessian_log_probit_likelihood(f, y, lik, single_precision=True):
    """probit/utilities.py:172-192."""
    ub, ub2, ub3 = BOUNDS["single" if single_precision else "double"]
    sigma = float(lik[0])
    Z, z1s, z2s = _safe_Z(f, y, lik, ub, ub2, ub3)
    p1 = _norm_pdf_inf(z1s)
    p2 = _norm_pdf_inf(z2s)
    w = grad_log_probit_likelihood(f, y, lik, single_precision)
    _z1 = np.where(np.isinf(z1s), 0.0, z1s)
    _z2 = np.where(np.isinf(z2s), 0.0, z2s)
    with np.errstate(all="ignore"):
        V = -(w**2) + (_z1 * p1 - _z2 * p2) / Z / sigma**2
    V = np.where(z1s > ub3, -(sigma**-2), V)
    V = np.where(z2s < -ub3, -(sigma**-2), V)
    return V


# ----------------------------------------------------------------------------
# Likelihood dispatch
# ----------------------------------------------------------------------------
GAUSSIAN = "gaussian"
PROBIT = "probit"
PROBIT_CLOSED = "probit_closed"  # BOUNDS["single"] (the callables' default)
PROBIT_CLOSED_DOUBLE = "probit_closed_double"  # single_precision=False: BOUNDS["double"]


def lik_derivs(kind, f, y, lik, order=2):
    if kind == GAUSSIAN:
        return gaussian_derivs(f, y, lik, order)
    if kind == PROBIT:
        return probit_derivs(f, y, lik, order)
    if kind in (PROBIT_CLOSED, PROBIT_CLOSED_DOUBLE):
        if order > 2:
            raise NotImplementedError("closed-form probit: value path only")
        sp = kind == PROBIT_CLOSED
        return {
            "ll": log_probit_likelihood(f, y, lik),
            "d1": grad_log_probit_likelihood(f, y, lik, single_precision=sp),
            "d2": hessian_log_probit_likelihood(f, y, lik, single_precision=sp),
        }
    raise ValueError(kind)


# ----------------------------------------------------------------------------
# Fixed-point drivers (probit/implicit/solvers.py)
# ----------------------------------------------------------------------------
def fixed_point_iteration(g, z0, tol, maxiter=100):
    """jaxopt.FixedPointIteration(fixed_point_fun=g, tol).run(z0) as called at
    probit/implicit/solvers.py:13-14: iterate z <- g(z) until
    ||g(z) - z||_2 <= tol (or maxiter), returning the last g(z).
    (jaxopt is absent; the stopping rule is restated from its documentation.)
    """
    z = np.array(z0, dtype=np.float64)
    err = np.inf
    it = 0
    while err > tol and it < maxiter:
        z_new = g(z)
        err = float(np.linalg.norm(z_new - z))
        z = z_new
        it += 1
    return z, it, err


def newton_map_lu(K, y, kind, lik):
    """g(z) = z - solve(jacobian(f_root)(z), f_root(z)) with
    f_root(z) = grad log p(K z) - z, dense Jacobian and LU solve, exactly as
    probit/implicit/solvers.py:18-25 with f = f_LA (probit/implicit/Laplace.py:4-9).
    """
    N = K.shape[0]

    def g(z):
        f = K @ z
        d = lik_derivs(kind, f, y, lik)
        J = d["d2"][:, None] * K - np.eye(N)
        return z - np.linalg.solve(J, d["d1"] - z)

    return g


def newton_map_bform(K, y, kind, lik):
    """The same Newton map through B = I + W^1/2 K W^1/2 (Rasmussen & Williams
    Alg. 3.1 identity (I + W K)^-1 = I - W^1/2 B^-1 W^1/2 K); needs W >= 0."""
    N = K.shape[0]

    def g(z):
        f = K @ z
        d = lik_derivs(kind, f, y, lik)
        W = -d["d2"]
        s = np.sqrt(W)
        B = np.eye(N) + s[:, None] * K * s[None, :]
        c = sla.cho_factor(B, lower=True)
        r = d["d1"] - z
        return z + r - s * sla.cho_solve(c, s * (K @ r))

    return g


def laplace_weight(K, y, kind, lik, tol=1e-5, maxiter=100, form="lu"):
    """LaplaceGP.weight (probit/approximators.py:265-269): Newton from zeros."""
    g = newton_map_lu(K, y, kind, lik) if form == "lu" else newton_map_bform(K, y, kind, lik)
    return fixed_point_iteration(g, np.zeros(K.shape[0]), tol, maxiter)


def laplace_weight_converged(K, y, kind, lik, maxiter=200):
    """Newton iterated to the FP64 fixed point (step <= 1e-12 or stagnation)."""
    g = newton_map_lu(K, y, kind, lik)
    z = np.zeros(K.shape[0])
    prev = np.inf
    for it in range(maxiter):
        z_new = g(z)
        err = float(np.linalg.norm(z_new - z))
        z = z_new
        if err <= 1e-12 * max(1.0, float(np.linalg.norm(z))) or (err >= prev and err < 1e-9):
            break
        prev = err
    return z


def laplace_precision(K, y, kind, lik, weight):
    """LaplaceGP.precision (probit/approximators.py:271-277): W = -hess(K w)."""
    f = K @ weight
    return -lik_derivs(kind, f, y, lik)["d2"], f


# ----------------------------------------------------------------------------
# Predict / objective (probit/approximators.py:154-197, probit/implicit/Laplace.py)
# ----------------------------------------------------------------------------
def predict(spec, X, Xt, weight, precision):
    """Approximator.predict (probit/approximators.py:154-180): latent mean and
    variance with K = Kff + diag(1/W) and a dense solve."""
    Kss = gram_elwise(spec, Xt)
    Kfs = gram(spec, X, Xt)
    K = gram(spec, X) + np.diag(1.0 / precision)
    var = Kss - np.einsum("ij,ij->j", Kfs, np.linalg.solve(K, Kfs))
    mean = Kfs.T @ weight
    return mean, var


def predict_covariance(spec, X, Xt, weight, precision):
    """Approximator.predict_covariance (probit/approximators.py:182-197)."""
    Kss = gram(spec, Xt, Xt)
    Kfs = gram(spec, X, Xt)
    K = gram(spec, X) + np.diag(1.0 / precision)
    return Kss - Kfs.T @ np.linalg.solve(K, Kfs)


def objective_LA(spec, X, y, kind, lik, weight):
    """objective_LA (probit/implicit/Laplace.py:12-30)."""
    K = gram(spec, X)
    f = K @ weight
    d = lik_derivs(kind, f, y, lik)
    W = -d["d2"]
    L = np.linalg.cholesky(K + np.diag(1.0 / W))
    return (-np.sum(d["ll"]) + 0.5 * f @ weight + np.sum(np.log(np.diag(L)))
            + 0.5 * np.sum(np.log(W)))


def laplace_objective_gradient(spec, X, y, kind, lik, weight):
    """Closed-form total derivative of objective_LA at the Laplace fixed point,
    including the implicit dependence of the weight on the parameters
    (the quantity the reference obtains from jax.value_and_grad through
    fixed_point_layer, probit/approximators.py:132-134, solvers.py:28-64).

    Derivation (Rasmussen & Williams Alg. 5.1 generalised to likelihood
    parameters): with C = W^1/2 B^-1 W^1/2, Sigma = K - K C K and
    s2 = -1/2 diag(Sigma) * d3,
      d/dtheta_K = -1/2 a'Kd a + 1/2 tr(C Kd) + s2' (I - K C) Kd a
      d/dtheta_l = -sum L_p + 1/2 sum diag(Sigma) (-L_ffp) + s2' (I - K C) K L_fp
    Returns dict name -> value; cutpoint grads are indexed by cutpoint slot.
    """
    K = gram(spec, X)
    N = K.shape[0]
    a = weight
    f = K @ a
    d = lik_derivs(kind, f, y, lik, order=3)
    W = -d["d2"]
    s = np.sqrt(W)
    B = np.eye(N) + s[:, None] * K * s[None, :]
    Binv = np.linalg.inv(B)
    C = s[:, None] * Binv * s[None, :]
    Sig_diag = np.diag(K) - np.einsum("ij,ji->i", K, C @ K)
    s2 = -0.5 * Sig_diag * d["d3"]
    s2hat = s2 - C @ (K @ s2)  # (I - C K) s2
    grads = {}
    for name, Kd in gram_param_derivs(spec, X):
        grads[name] = (-0.5 * a @ Kd @ a + 0.5 * np.sum(C * Kd) + s2hat @ (Kd @ a))

    def lik_term(Lp, Lfp, Lffp, mask=None):
        m = np.ones(N) if mask is None else mask
        return (-np.sum(m * Lp) + 0.5 * np.sum(m * Sig_diag * (-Lffp))
                + s2hat @ (K @ (m * Lfp)))

    Lp, Lfp, Lffp = d["sigma"]
    grads["noise_std"] = lik_term(Lp, Lfp, Lffp)
    if kind == PROBIT:
        cut = np.asarray(lik[1], dtype=np.float64)
        yi = np.asarray(y).astype(np.int64)
        g = np.zeros(cut.shape[0])
        for j in range(cut.shape[0]):
            if np.isinf(cut[j]):
                continue
            lo = (yi == j).astype(np.float64)  # cutpoint j is the lower end
            hi = (yi + 1 == j).astype(np.float64)  # cutpoint j is the upper end
            g[j] = lik_term(*d["b_lo"], mask=lo) + lik_term(*d["b_hi"], mask=hi)
        grads["cutpoints"] = g
    return grads


# ----------------------------------------------------------------------------
# VB (probit/implicit/VB.py, probit/approximators.py:280-339)
# ----------------------------------------------------------------------------
def f_VB(K, y, kind, lik, weight):
    """probit/implicit/VB.py:4-16."""
    N = K.shape[0]
    sigma = float(lik[0])
    f = K @ weight
    c = sla.cho_factor(sigma**2 * np.eye(N) + K, lower=True)
    return sla.cho_solve(c, f + sigma * lik_derivs(kind, f, y, lik)["d1"])


def vb_weight(K, y, kind, lik, tol=1e-5, maxiter=100):
    """VBGP.weight (probit/approximators.py:332-334): plain FPI from zeros."""
    return fixed_point_iteration(lambda w: f_VB(K, y, kind, lik, w), np.zeros(K.shape[0]),
                                 tol, maxiter)


def vb_precision(lik, N):
    """VBGP.precision (probit/approximators.py:336-339)."""
    return np.full(N, 1.0 / float(lik[0]) ** 2)


def objective_VB(spec, X, y, kind, lik, weight):
    """probit/implicit/VB.py:19-40."""
    K = gram(spec, X)
    f = K @ weight
    N = X.shape[0]
    sigma = float(lik[0])
    L = np.linalg.cholesky(sigma**2 * np.eye(N) + K)
    LinvT = sla.solve_triangular(L, np.eye(N), lower=True)
    cov = sla.solve_triangular(L.T, LinvT, lower=False)
    log_det_cov = -2.0 * np.sum(np.log(np.diag(L)))
    trace_cov = np.sum(np.diag(cov))
    trace_post = np.sum(K * cov)
    return (0.5 * trace_post + 0.5 * sigma**2 * trace_cov + 0.5 * f @ weight
            - N * math.log(sigma) - 0.5 * log_det_cov - 0.5 * N
            - np.sum(lik_derivs(kind, f, y, lik)["ll"]))


# ----------------------------------------------------------------------------
# Predictive class probabilities (probit/utilities.py:232-249)
# ----------------------------------------------------------------------------
def probit_predictive_distributions(lik, mean, variance):
    sigma, cut = float(lik[0]), np.asarray(lik[1], dtype=np.float64)
    J = cut.shape[0] - 1
    std = np.sqrt(variance + sigma**2)
    out = np.ones((mean.shape[0], J))
    for j in range(J):
        lo = 0.0 if cut[j] == -np.inf else ndtr((cut[j] - mean) / std)
        hi = 1.0 if cut[j + 1] == np.inf else ndtr((cut[j + 1] - mean) / std)
        out[:, j] = hi - lo
    return out


def calculate_metrics(y_test, predictive_distributions):
    """examples/classification.py:325-341: argmax predictions and the reference's
    metrics, with its column indexing predictive_distributions[:, y_test]
    (an M x M matrix whose log is summed).  Returns (mean_zero_one,
    mean_absolute_error, log_predictive_probability, predictive_likelihood,
    y_pred)."""
    P = np.asarray(predictive_distributions, dtype=np.float64)
    y = np.asarray(y_test).astype(np.int64)
    y_pred = np.argmax(P, axis=1)
    predictive_likelihood = P[:, y]
    mean_absolute_error = np.sum(np.abs(y_pred - y)) / len(y)
    with np.errstate(divide="ignore"):
        log_predictive_probability = np.sum(np.log(predictive_likelihood))
    mean_zero_one = np.sum(y_pred != y) / len(y)
    return mean_zero_one, mean_absolute_error, log_predictive_probability, \
        predictive_likelihood, y_pred


# ----------------------------------------------------------------------------
# GP prior samples (examples/regression.py:52-58, examples/classification.py:227-234)
# ----------------------------------------------------------------------------
def prior_sample(spec: KernelSpec, X: np.ndarray, z: np.ndarray, jitter: float = 1e-10):
    """f = L_K z with L_K = chol(K + jitter I), K = kernel(X) (the generators'
    ``K = B.dense(kernel(X)); K = K + jitter * I; L_K = cholesky(K); f = L_K @ z``).
    z is (N,) or (N, S); f has its shape."""
    K = gram(spec, X) + jitter * np.eye(X.shape[0])
    L = np.linalg.cholesky(K)
    return L @ z
