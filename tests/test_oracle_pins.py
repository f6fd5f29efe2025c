"""Pins of the two oracle pieces that no library routine checks directly:
the fixed-point loop driver (jaxopt.FixedPointIteration as called at
probit/implicit/solvers.py:13-14) and the branch selection of the
closed-form probit likelihood's class probability (_safe_Z,
probit/utilities.py:88-148).

Fixed-point loop: on a linear map g(z) = rho z + c from z0 = 0 the iterates
are z_n = c (1 - rho^n) / (1 - rho) and the step norms ||z_n - z_{n-1}|| =
|rho|^(n-1) ||c|| in closed form, so the iteration count, the returned iterate
(g of the last input, not the input), the <= comparison with an absolute L2
tolerance and the default cap of 100 iterations are all fixed analytically.
Variants with one of those choices changed are run through the same checks
and must fail them.

_safe_Z: every regime is checked against SciPy's normal CDF.  The tail
approximation T(z) = phi(z)/z exp(h(z)) underestimates Q(z) = 1 - Phi(z) for
z >= 1.3 by a relative amount below 100/z^8 (the next term of the asymptotic
Mills-ratio expansion behind h has coefficient 88.25); exact regimes must equal
the CDF difference; the linear regimes return the placeholder 1 with E = z/sigma,
V = -1/sigma^2; and the lower branches are the exact mirror image of the upper
ones (T is odd), so a slip in one branch breaks the symmetry (equal up to
rounding).  With the double
bounds [2.3, 3.6, 4.8] Z is positive everywhere; with the reference's default
single bounds [1.3, 1.8, 2.3] the two-sided tail gives a negative Z for narrow
class intervals with z1 just above 1.3 (T is not monotone there): a property
of the reference's own arithmetic, recorded here so that it stays reproduced.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.special import ndtr

from oracle import gp_oracle as O


# ------------------------------------------------------------- fixed-point loop ---
def _loop_matches_closed_form(fpi) -> bool:
    """True if `fpi(g, z0, tol, maxiter)` behaves as the jaxopt loop on linear
    maps whose iterates are known in closed form."""
    ok = True
    # rho = 1/2, ||c|| = 1, tol = 2^-10: step n has norm exactly 2^-(n-1), so
    # the first step with norm <= tol is n = 11 (n = 12 under a strict <)
    c = np.zeros(5)
    c[2] = 1.0
    z, it, err = fpi(lambda v: 0.5 * v + c, np.zeros(5), 2.0 ** -10, 100)
    ok &= it == 11 and err == 2.0 ** -10
    ok &= bool(np.array_equal(z, c * (1.0 - 0.5 ** 11) / 0.5))
    # a large offset: an absolute tolerance stops as above, a relative one
    # (step / ||z||) would stop much earlier
    c2 = 1e6 * c
    z, it, err = fpi(lambda v: 0.5 * v + c2, np.zeros(5), 1e6 * 2.0 ** -10, 100)
    ok &= it == 11
    z, it, err = fpi(lambda v: 0.5 * v + c2, np.zeros(5), 2.0 ** -10, 100)
    ok &= it == 31  # 1e6 * 2^-(n-1) <= 2^-10 first at n - 1 = 30 (2^20 > 1e6)
    # slow contraction: the cap of 100 iterations ends the loop
    rho = np.array([0.999, -0.5, 0.25])
    c3 = np.array([1.0, 2.0, -3.0])
    z, it, err = fpi(lambda v: rho * v + c3, np.zeros(3), 1e-12, 100)
    ok &= it == 100
    ok &= bool(np.allclose(err, np.linalg.norm(rho ** 99 * c3), rtol=1e-12))
    ok &= bool(np.allclose(z, c3 * (1.0 - rho ** 100) / (1.0 - rho), rtol=1e-12))
    # mixed rates, tolerance between two step norms: the count is the first n
    # whose closed-form step norm is <= tol
    tol = 3e-4
    rho2 = np.array([0.9, -0.5, 0.25])
    norms = [np.linalg.norm(rho2 ** (n - 1) * c3) for n in range(1, 500)]
    n_expect = 1 + next(i for i, v in enumerate(norms) if v <= tol)
    z, it, err = fpi(lambda v: rho2 * v + c3, np.zeros(3), tol, 100)
    ok &= it == n_expect
    ok &= bool(np.allclose(z, c3 * (1.0 - rho2 ** it) / (1.0 - rho2), rtol=1e-12))
    # z0 = the fixed point: one iteration, a zero step
    zs = c3 / (1.0 - rho)
    z, it, err = fpi(lambda v: rho * v + c3, zs, 1e-12, 100)
    ok &= it == 1 and err <= 1e-15
    return bool(ok)


def test_fixed_point_loop_matches_closed_form_iterates():
    assert _loop_matches_closed_form(O.fixed_point_iteration)


def _variant(strict=False, return_input=False, relative=False, cap=None):
    def fpi(g, z0, tol, maxiter=100):
        z = np.array(z0, dtype=np.float64)
        err, it, last_in = np.inf, 0, z
        cap_ = maxiter if cap is None else cap
        while (err >= tol if strict else err > tol) and it < cap_:
            z_new = g(z)
            err = float(np.linalg.norm(z_new - z))
            if relative:
                err /= max(1e-300, float(np.linalg.norm(z_new)))
            last_in, z = z, z_new
            it += 1
        return (last_in if return_input else z), it, err

    return fpi


@pytest.mark.parametrize("mutant", [
    _variant(strict=True),          # stops on err < tol instead of err <= tol
    _variant(return_input=True),    # returns z_k instead of g(z_k)
    _variant(relative=True),        # relative instead of absolute tolerance
    _variant(cap=99),               # another iteration cap
])
def test_fixed_point_loop_check_rejects_plausible_mistakes(mutant):
    assert not _loop_matches_closed_form(mutant)


# ------------------------------------------------------------------- _safe_Z ---
DOUBLE = tuple(O.BOUNDS["double"])
SINGLE = tuple(O.BOUNDS["single"])


def _Z(z1, z2, bounds, sigma=0.7, shift=0.3):
    """_safe_Z at the z-scores (z1, z2) (either may be infinite): the class
    interval [shift + sigma z1, shift + sigma z2] around f = shift, class 1 of
    three (two-sided), or the bottom / top class for an infinite end."""
    if math.isinf(z1) and z1 < 0:
        cut = np.array([-np.inf, shift + sigma * z2, shift + sigma * z2 + 1.0, np.inf])
        y = 0
    elif math.isinf(z2) and z2 > 0:
        cut = np.array([-np.inf, shift + sigma * z1 - 1.0, shift + sigma * z1, np.inf])
        y = 2
    else:
        cut = np.array([-np.inf, shift + sigma * z1, shift + sigma * z2, np.inf])
        y = 1
    f, yv, lp = np.array([shift]), np.array([y]), (sigma, cut)
    Z, z1s, z2s = O._safe_Z(f, yv, lp, *bounds)
    E = O.grad_log_probit_likelihood(f, yv, lp, single_precision=bounds == SINGLE)
    V = O.hessian_log_probit_likelihood(f, yv, lp, single_precision=bounds == SINGLE)
    return float(Z[0]), float(z1s[0]), float(z2s[0]), float(E[0]), float(V[0])


def _Q(z):
    return float(ndtr(-z)) if math.isfinite(z) else (0.0 if z > 0 else 1.0)


def _eps(z):
    """Bound on the relative error of the tail series at |z| >= 1.3."""
    return 0.0 if math.isinf(z) else 100.0 / abs(z) ** 8


def _regime_points(bounds):
    ub, ub2, ub3 = bounds
    pts = {}
    pts["exact"] = [(-ub, ub), (-0.4, 0.9), (ub - 0.5, ub), (-math.inf, -ub), (ub, math.inf),
                    (-1.0, 1.0)]
    lo1 = np.linspace(ub + 1e-3, ub2, 5)
    pts["two_sided_upper"] = [(a, a + d) for a in lo1 for d in (0.3, 1.0)] + \
        [(a, math.inf) for a in lo1]
    lo2 = np.linspace(ub2 + 1e-3, ub3 - 1e-3, 5)
    pts["one_sided_upper"] = [(a, a + d) for a in lo2 for d in (0.2, 1.0)] + \
        [(a, math.inf) for a in lo2]
    pts["linear_upper"] = [(ub3 + 1e-3, ub3 + 0.5), (ub3 + 2.0, math.inf), (9.0, 10.0)]
    return pts


@pytest.mark.parametrize("bounds", [DOUBLE, SINGLE], ids=["double", "single"])
def test_safe_Z_exact_regime_is_the_cdf_difference(bounds):
    for z1, z2 in _regime_points(bounds)["exact"]:
        Z, z1s, z2s, _, _ = _Z(z1, z2, bounds)
        ref = float(ndtr(z2s) - ndtr(z1s))
        assert abs(Z - ref) <= 1e-15 + 1e-13 * ref, (z1, z2, Z, ref)


@pytest.mark.parametrize("bounds", [DOUBLE, SINGLE], ids=["double", "single"])
def test_safe_Z_lower_branches_mirror_upper_branches(bounds):
    """Z(z1, z2) in an upper regime equals Z(-z2, -z1) in the mirrored lower
    regime up to rounding (h is even, so the tail term is odd)."""
    pts = _regime_points(bounds)
    for key in ("two_sided_upper", "one_sided_upper", "linear_upper", "exact"):
        for z1, z2 in pts[key]:
            up = _Z(z1, z2, bounds, sigma=1.0, shift=0.0)
            lo = _Z(-z2, -z1, bounds, sigma=1.0, shift=0.0)
            # the same up to rounding (powers of -z, the CDF difference's order)
            assert lo[0] == pytest.approx(up[0], rel=1e-13, abs=1e-16), (key, z1, z2)
            if key == "linear_upper" or key.startswith("two_sided"):
                # E flips sign under the mirror; V does not
                assert -lo[3] == pytest.approx(up[3], rel=1e-12), (key, z1, z2)
                assert lo[4] == pytest.approx(up[4], rel=1e-12), (key, z1, z2)


def test_safe_Z_tail_regimes_against_scipy_double_bounds():
    """Double bounds: positive everywhere and within the series accuracy of
    SciPy's CDF difference (two-sided tails) or upper / lower tail mass
    (one-sided tails); lower branches checked directly too."""
    pts = _regime_points(DOUBLE)
    for z1, z2 in pts["two_sided_upper"]:
        Z, z1s, z2s, E, _ = _Z(z1, z2, DOUBLE)
        ref = _Q(z1s) - _Q(z2s)
        assert Z > 0.0, (z1, z2, Z)
        assert abs(Z - ref) <= _eps(z1s) * _Q(z1s) + _eps(z2s) * _Q(z2s), (z1, z2, Z, ref)
        # E = (phi(z1) - phi(z2)) / (sigma Z): the exact derivative of log(dPhi)
        # to the same relative accuracy
        e_ref = O.probit_derivs(np.array([0.3]), np.array([1 if math.isfinite(z2) else 2]),
                                (0.7, _cuts(z1, z2)))["d1"][0]
        tolr = (_eps(z1s) * _Q(z1s) + _eps(z2s) * _Q(z2s)) / ref
        assert abs(E - e_ref) <= 1.05 * tolr / (1 - tolr) * abs(e_ref) + 1e-9, (z1, z2, E, e_ref)
    for z1, z2 in pts["one_sided_upper"]:
        Z, z1s, z2s, _, _ = _Z(z1, z2, DOUBLE)
        assert Z > 0.0
        # the upper tail beyond z1, whatever z2: a two-sided formula here would
        # miss it by Q(z2), far outside the bound for the narrow intervals
        assert abs(Z - _Q(z1s)) <= _eps(z1s) * _Q(z1s), (z1, z2, Z, _Q(z1s))
    for z1, z2 in pts["two_sided_upper"]:  # lower two-sided: -z2 <= z < -z1
        Z, z1s, z2s, _, _ = _Z(-z2, -z1, DOUBLE)
        ref = float(ndtr(z2s)) - (float(ndtr(z1s)) if math.isfinite(z1s) else 0.0)
        assert Z > 0.0
        assert abs(Z - ref) <= _eps(z1s) * (1.0 - _Q(z1s) if math.isfinite(z1s) else 0.0) + \
            _eps(z2s) * float(ndtr(z2s)), (z1, z2, Z, ref)
    for z1, z2 in pts["one_sided_upper"]:  # lower one-sided: the lower tail below z2
        Z, z1s, z2s, _, _ = _Z(-z2, -z1, DOUBLE)
        assert Z > 0.0
        assert abs(Z - float(ndtr(z2s))) <= _eps(z2s) * float(ndtr(z2s)), (z1, z2)


def _cuts(z1, z2, sigma=0.7, shift=0.3):
    if math.isinf(z2):
        return np.array([-np.inf, shift + sigma * z1 - 1.0, shift + sigma * z1, np.inf])
    return np.array([-np.inf, shift + sigma * z1, shift + sigma * z2, np.inf])


@pytest.mark.parametrize("bounds", [DOUBLE, SINGLE], ids=["double", "single"])
def test_safe_Z_linear_regimes(bounds):
    """Beyond the third bound Z is the placeholder 1, E = z / sigma and
    V = -1 / sigma^2 (probit/utilities.py:142-148, 166-167, 190-191)."""
    sigma = 0.7
    for z1, z2 in _regime_points(bounds)["linear_upper"]:
        Z, z1s, _, E, V = _Z(z1, z2, bounds, sigma=sigma)
        assert Z == 1.0
        assert E == pytest.approx(z1s / sigma, rel=1e-15)
        assert V == pytest.approx(-1.0 / sigma ** 2, rel=1e-15)
        Z, _, z2s, E, V = _Z(-z2, -z1, bounds, sigma=sigma)
        assert Z == 1.0
        assert E == pytest.approx(z2s / sigma, rel=1e-15)
        assert V == pytest.approx(-1.0 / sigma ** 2, rel=1e-15)


def test_safe_Z_single_bounds_accuracy_and_reference_quirk():
    """Single bounds: within the (looser, small-z) series bound of SciPy
    everywhere the tail is used, positive where the class interval is wide,
    and negative for a narrow interval just above the first bound -- the
    reference's arithmetic, kept as it is."""
    pts = _regime_points(SINGLE)
    for z1, z2 in pts["two_sided_upper"] + pts["one_sided_upper"]:
        Z, z1s, z2s, _, _ = _Z(z1, z2, SINGLE)
        two = (z1, z2) in pts["two_sided_upper"]
        ref = _Q(z1s) - _Q(z2s) if two else _Q(z1s)
        bound = _eps(z1s) * _Q(z1s) + (_eps(z2s) * _Q(z2s) if two else 0.0)
        assert abs(Z - ref) <= bound, (z1, z2, Z, ref)
    for z1 in (1.5, 1.8):
        assert _Z(z1, math.inf, SINGLE)[0] > 0.0  # wide: the top class
    Z = _Z(1.35, 1.6, SINGLE, sigma=1.0, shift=0.0)[0]
    assert Z < 0.0 and Z == pytest.approx(-0.013942727, rel=1e-6)


@pytest.mark.parametrize("mutation", ["unsigned_tail", "negated_tail"])
def test_safe_Z_checks_reject_plausible_mistakes(monkeypatch, mutation):
    """A tail term that loses its sign for negative z, or has the wrong sign
    overall, fails the checks above."""
    orig = O._Z_far_tails
    if mutation == "unsigned_tail":
        monkeypatch.setattr(O, "_Z_far_tails", lambda z: orig(np.abs(z)))
    else:
        monkeypatch.setattr(O, "_Z_far_tails", lambda z: -orig(z))
    with pytest.raises(AssertionError):
        test_safe_Z_tail_regimes_against_scipy_double_bounds()
