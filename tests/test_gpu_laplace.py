"""Parity of the drop-in LaplaceGP / VBGP / serving path against the oracle,
at sizes the oracle finishes in seconds."""

import numpy as np
import pytest

from oracle import gp_oracle as O
from tests.helpers import lik_kind, make_prior, rel_err, small_problem

pytestmark = pytest.mark.gpu


def _gp(cls_name, X, y, kind, lik, tol=1e-5, **kw):
    import probit_amd as P
    cls = getattr(P, cls_name)
    ll = P.log_gaussian_likelihood if lik == "gaussian" else P.log_probit_likelihood
    return cls(data=(X, y), prior=make_prior(kind), log_likelihood=ll, tolerance=tol, **kw)


@pytest.mark.parametrize("lik", ["gaussian", "probit"])
@pytest.mark.parametrize("kind", ["EQ", "Matern12"])
@pytest.mark.parametrize("N,D", [(37, 1), (200, 3), (1100, 2)])
def test_laplace_matches_reference_iterates(lik, kind, N, D):
    X, y, theta, lp, spec = small_problem(N, N, D, lik, kind)
    gp = _gp("LaplaceGP", X, y, kind, lik)
    params = (theta, lp)
    w, p = gp.approximate_posterior(params)
    K = O.gram(spec, X)
    w_ref, it_ref, _ = O.laplace_weight(K, y, lik_kind(lik), lp, tol=1e-5)
    assert gp.last_result.iterations == it_ref
    assert rel_err(w.cpu().numpy(), w_ref) < 1e-9
    p_ref, _ = O.laplace_precision(K, y, lik_kind(lik), lp, w_ref)
    assert rel_err(p.cpu().numpy(), p_ref) < 1e-9
    Xt = np.random.default_rng(5).normal(size=(150, D))
    mean, var = gp.predict(Xt, params, w, p)
    m_ref, v_ref = O.predict(spec, X, Xt, w_ref, p_ref)
    assert rel_err(mean.cpu().numpy(), m_ref) < 1e-9
    assert rel_err(var.cpu().numpy(), v_ref) < 1e-9
    obj = float(gp.objective()(params))
    ref = O.objective_LA(spec, X, y, lik_kind(lik), lp, w_ref)
    assert abs(obj - ref) <= 1e-9 * abs(ref)


@pytest.mark.parametrize("lik", ["gaussian", "probit"])
@pytest.mark.parametrize("kind,period", [("EQ", None), ("Matern12", None), ("EQ", 0.9)])
def test_laplace_gradient_matches_closed_form(lik, kind, period):
    import probit_amd as P
    X, y, theta, lp, spec = small_problem(3, 120, 2, lik, kind, period=period)
    ll = P.log_gaussian_likelihood if lik == "gaussian" else P.log_probit_likelihood
    gp = P.LaplaceGP(data=(X, y), prior=make_prior(kind, period), log_likelihood=ll,
                     tolerance=1e-11)
    val, grads = gp.value_and_grad()((theta, lp))
    K = O.gram(spec, X)
    w = O.laplace_weight_converged(K, y, lik_kind(lik), lp)
    ref = O.laplace_objective_gradient(spec, X, y, lik_kind(lik), lp, w)
    assert abs(val - O.objective_LA(spec, X, y, lik_kind(lik), lp, w)) < 1e-9 * abs(val)
    (g_ls, g_var), g_lik = grads
    ref_ls = sum(ref[f"lengthscale{d}"] for d in range(2))
    assert abs(g_ls - ref_ls) <= 1e-7 * max(1.0, abs(ref_ls))
    assert abs(g_var - ref["variance"]) <= 1e-7 * max(1.0, abs(ref["variance"]))
    assert abs(g_lik[0] - ref["noise_std"]) <= 1e-7 * max(1.0, abs(ref["noise_std"]))
    if lik == "probit":
        assert np.allclose(g_lik[1], ref["cutpoints"], rtol=1e-7, atol=1e-9)


def test_periodic_gradient_reaches_period_parameter():
    """A prior whose period is a parameter gets d/dperiod from the adjoint."""
    import torch

    import probit_amd as P
    X, y, theta, lp, spec = small_problem(4, 90, 1, "gaussian", "EQ", period=0.7)

    def prior(th):
        ls, var, per = th
        return var * P.EQ().stretch(ls).periodic(per)

    gp = P.LaplaceGP(data=(X, y), prior=prior, log_likelihood=P.log_gaussian_likelihood,
                     tolerance=1e-11)
    val, grads = gp.value_and_grad()(((theta[0], theta[1], 0.7), lp))
    K = O.gram(spec, X)
    w = O.laplace_weight_converged(K, y, O.GAUSSIAN, lp)
    ref = O.laplace_objective_gradient(spec, X, y, O.GAUSSIAN, lp, w)
    assert abs(grads[0][2] - ref["period0"]) <= 1e-7 * max(1.0, abs(ref["period0"]))
    del torch


@pytest.mark.parametrize("lik", ["gaussian", "probit"])
def test_vb_matches_reference(lik):
    X, y, theta, lp, spec = small_problem(11, 96, 2, lik, "EQ", noise=1.0 if lik == "gaussian"
                                          else 0.8)
    gp = _gp("VBGP", X, y, "EQ", lik, tol=1e-5)
    params = (theta, lp)
    w, p = gp.approximate_posterior(params)
    K = O.gram(spec, X)
    w_ref, it_ref, _ = O.vb_weight(K, y, lik_kind(lik), lp, tol=1e-5)
    assert it_ref < 100 and gp.last_result.status == 0  # converged, in the contracting regime
    assert gp.last_result.iterations == it_ref
    assert rel_err(w.cpu().numpy(), w_ref) < 1e-8
    assert np.allclose(p.cpu().numpy(), O.vb_precision(lp, 96))
    obj = float(gp.objective()(params))
    ref = O.objective_VB(spec, X, y, lik_kind(lik), lp, w_ref)
    assert abs(obj - ref) <= 1e-8 * abs(ref)
    Xt = np.random.default_rng(2).normal(size=(64, 2))
    mean, var = gp.predict(Xt, params, w, p)
    m_ref, v_ref = O.predict(spec, X, Xt, w_ref, O.vb_precision(lp, 96))
    assert rel_err(mean.cpu().numpy(), m_ref) < 1e-8
    assert rel_err(var.cpu().numpy(), v_ref) < 1e-8


def test_predict_with_foreign_weight_and_precision():
    """predict() with arrays that did not come from this GP re-factors from them."""
    X, y, theta, lp, spec = small_problem(12, 80, 2, "probit", "EQ")
    gp = _gp("LaplaceGP", X, y, "EQ", "probit")
    K = O.gram(spec, X)
    w = O.laplace_weight_converged(K, y, O.PROBIT, lp)
    p, _ = O.laplace_precision(K, y, O.PROBIT, lp, w)
    Xt = np.random.default_rng(3).normal(size=(40, 2))
    mean, var = gp.predict(Xt, (theta, lp), w, p)
    m_ref, v_ref = O.predict(spec, X, Xt, w, p)
    assert rel_err(mean.cpu().numpy(), m_ref) < 1e-10
    assert rel_err(var.cpu().numpy(), v_ref) < 1e-10


@pytest.mark.parametrize("cls,lik,mode,M", [("LaplaceGP", "probit", "fp64", 100),
                                            ("LaplaceGP", "gaussian", "fp64", 64),
                                            ("LaplaceGP", "probit", "fp32", 130),
                                            ("VBGP", "probit", "fp64", 70)])
def test_predict_covariance_matches_reference(cls, lik, mode, M):
    """predict_covariance (probit/approximators.py:182-197) against the oracle's
    dense solve; its diagonal against predict's variance; exact symmetry."""
    # VB: sigma >= 0.5, where its fixed-point iteration contracts
    X, y, theta, lp, spec = small_problem(21, 150, 2, lik, "Matern12",
                                          noise=0.6 if cls == "VBGP" else 0.3)
    gp = _gp(cls, X, y, "Matern12", lik, precision_mode=mode)
    params = (theta, lp)
    w, p = gp.approximate_posterior(params)
    assert not gp.last_result.status & 4, "fixed point not reached"  # INFO_MAXITER
    Xt = np.random.default_rng(22).normal(size=(M, 2))
    cov = gp.predict_covariance(Xt, params, w, p).cpu().numpy()
    _, var = gp.predict(Xt, params, w, p)
    ref = O.predict_covariance(spec, X, Xt, w.cpu().numpy(), p.cpu().numpy())
    assert cov.shape == (M, M)
    assert np.array_equal(cov, cov.T)
    tol = 1e-10 if mode == "fp64" else 1e-5
    assert np.max(np.abs(cov - ref)) <= tol * np.max(np.abs(ref))
    # FP64: the diagonal is predict's variance.  FP32: predict's variance gets
    # the FP64 correction of the quadratic form, the covariance keeps the FP32
    # solve (its bound above), so they agree at that bound, and the variance
    # meets a much tighter one against the oracle
    v = var.cpu().numpy()
    assert np.max(np.abs(np.diag(cov) - v)) <= (1e-12 if mode == "fp64" else tol) * float(theta[1])
    v_ref = np.diag(ref)
    assert np.max(np.abs(v - v_ref)) <= (1e-10 if mode == "fp64" else 1e-9) * np.max(v_ref)


def test_precision_and_construct():
    X, y, theta, lp, spec = small_problem(13, 70, 2, "probit", "Matern12")
    gp = _gp("LaplaceGP", X, y, "Matern12", "probit")
    K = O.gram(spec, X)
    w = O.laplace_weight_converged(K, y, O.PROBIT, lp)
    W, f = gp.precision(w, (theta, lp))
    W_ref, f_ref = O.laplace_precision(K, y, O.PROBIT, lp, w)
    assert rel_err(W.cpu().numpy(), W_ref) < 1e-11
    assert rel_err(f.cpu().numpy(), f_ref) < 1e-11
    g = gp.construct()((theta, lp), w)
    g_ref = O.lik_derivs(O.PROBIT, f_ref, y, lp)["d1"]
    assert rel_err(g.cpu().numpy(), g_ref) < 1e-10


@pytest.mark.parametrize("mode,tol", [("fp32", 1e-5), ("bf16", 1e-5)])
def test_mixed_precision_mean_within_tolerance(mode, tol):
    """Mixed-precision factors with FP64 residuals and refinement keep the
    predictive mean within 1e-5 relative of the FP64 oracle."""
    X, y, theta, lp, spec = small_problem(21, 700, 3, "probit", "EQ", ls=1.0, var=1.0,
                                          noise=0.5)
    gp = _gp("LaplaceGP", X, y, "EQ", "probit", precision_mode=mode)
    params = (theta, lp)
    w, p = gp.approximate_posterior(params)
    Xt = np.random.default_rng(4).normal(size=(256, 3))
    mean, var = gp.predict(Xt, params, w, p)
    K = O.gram(spec, X)
    w_ref = O.laplace_weight_converged(K, y, O.PROBIT, lp)
    p_ref, _ = O.laplace_precision(K, y, O.PROBIT, lp, w_ref)
    m_ref, v_ref = O.predict(spec, X, Xt, w_ref, p_ref)
    assert rel_err(mean.cpu().numpy(), m_ref) < tol


def test_serving_alone_equals_in_mix():
    """A request's results are bit-identical alone and in a mixed batch, and
    match the single-problem entry point and the oracle."""
    from probit_amd.serving import ServingEngine
    probs = []
    for i, (N, D, lik, kind) in enumerate([(150, 2, "probit", "EQ"), (300, 1, "gaussian",
                                                                      "Matern12"),
                                           (64, 4, "probit", "Matern12"),
                                           (500, 3, "gaussian", "EQ")]):
        X, y, theta, lp, spec = small_problem(30 + i, N, D, lik, kind)
        Xt = np.random.default_rng(40 + i).normal(size=(100, D))
        probs.append((X, y, theta, lp, spec, lik, kind, Xt))

    def run(idx, mode="fp64"):
        eng = ServingEngine(precision=mode, capacity=8, hbm_budget_bytes=1 << 32)
        for i in idx:
            X, y, theta, lp, spec, lik, kind, Xt = probs[i]
            assert eng.submit(i, X, y, kind, theta[0], theta[1], lik, lp[0],
                              lp[1] if lik == "probit" else None, Xt)
        out = {}
        while eng.in_flight:
            for c in eng.step():
                out[c.tag] = c
        eng.close()
        return out

    alone = run([0])
    mix = run([3, 1, 0, 2])
    for k in ("mean", "var", "weight", "precision"):
        a = getattr(alone[0], k).cpu().numpy()
        b = getattr(mix[0], k).cpu().numpy()
        assert np.array_equal(a, b), k
    for i in range(4):
        X, y, theta, lp, spec, lik, kind, Xt = probs[i]
        K = O.gram(spec, X)
        w_ref, it_ref, _ = O.laplace_weight(K, y, lik_kind(lik), lp, tol=1e-5)
        p_ref, _ = O.laplace_precision(K, y, lik_kind(lik), lp, w_ref)
        m_ref, v_ref = O.predict(spec, X, Xt, w_ref, p_ref)
        c = mix[i]
        assert c.iterations == it_ref
        assert rel_err(c.mean.cpu().numpy(), m_ref) < 1e-9
        assert rel_err(c.var.cpu().numpy(), v_ref) < 1e-9
    # same in mixed precision: alone == in the mix, bit for bit
    a = run([2], "bf16")
    b = run([0, 1, 2, 3], "bf16")
    assert np.array_equal(a[2].mean.cpu().numpy(), b[2].mean.cpu().numpy())
    assert np.array_equal(a[2].var.cpu().numpy(), b[2].var.cpu().numpy())


@pytest.mark.parametrize("mode,n", [("fp64", 1024), ("fp32", 1024), ("fp32", 1088)])
def test_serving_schedule_independent(mode, n):
    """A member's results do not depend on the factorisation schedule its
    batch gets: alone, the N = 1024 request is factored by the one-launch-per-
    column kernel throughout; next to fifteen N = 2560 requests its first
    super-steps go through the separate update / factor / panel kernels
    (N = 1088: an odd number of block columns, a partial last super-step)."""
    from probit_amd.serving import ServingEngine
    target = small_problem(90, n, 2, "probit", "EQ")
    Xt = np.random.default_rng(91).normal(size=(64, 2))
    others = [small_problem(100 + i, 2560, 2, "gaussian", "Matern12") for i in range(15)]

    def run(with_others):
        eng = ServingEngine(precision=mode, capacity=16, hbm_budget_bytes=1 << 34)
        X, y, theta, lp, _ = target
        assert eng.submit(0, X, y, "EQ", theta[0], theta[1], "probit", lp[0], lp[1], Xt)
        if with_others:
            for i, (Xo, yo, th, lpo, _) in enumerate(others):
                assert eng.submit(1 + i, Xo, yo, "Matern12", th[0], th[1], "gaussian", lpo[0],
                                  None, Xt)
        out = {}
        while eng.in_flight:
            for c in eng.step():
                out[c.tag] = c
        eng.close()
        return out[0]

    a, b = run(False), run(True)
    assert a.iterations == b.iterations
    for k in ("mean", "var", "weight", "precision"):
        assert np.array_equal(getattr(a, k).cpu().numpy(), getattr(b, k).cpu().numpy()), k


def test_serving_bf16_escalates_like_single_fit():
    """A BF16 request whose factor loses definiteness (ill-conditioned B:
    smooth prior, large signal variance, small probit noise) is redone with
    FP32 updates, flagged PGP_INFO_ESCALATED, and matches the single-problem
    fit, which escalates the same way, bit for bit; well-conditioned
    requests beside it are unaffected."""
    import probit_amd as P
    from probit_amd.serving import ServingEngine
    bad = small_problem(95, 640, 1, "probit", "EQ", ls=2.0, var=10.0, noise=0.2)
    good = [small_problem(96 + i, 200, 2, "probit", "Matern12") for i in range(2)]
    Xt = np.random.default_rng(97).normal(size=(50, 1))
    eng = ServingEngine(precision="bf16", capacity=4, hbm_budget_bytes=1 << 32)
    X, y, theta, lp, spec = bad
    assert eng.submit(0, X, y, "EQ", theta[0], theta[1], "probit", lp[0], lp[1], Xt)
    for i, (Xg, yg, th, lpg, _) in enumerate(good):
        Xtg = np.random.default_rng(98 + i).normal(size=(50, 2))
        assert eng.submit(1 + i, Xg, yg, "Matern12", th[0], th[1], "probit", lpg[0], lpg[1], Xtg)
    out = {}
    while eng.in_flight:
        for c in eng.step():
            out[c.tag] = c
    eng.close()
    gp = _gp("LaplaceGP", X, y, "EQ", "probit", precision_mode="bf16")
    w, p = gp.approximate_posterior((theta, lp))
    mean, var = gp.predict(Xt, (theta, lp), w, p)
    assert gp.last_result.status & P._lib.INFO_ESCALATED
    c = out[0]
    assert c.status & P._lib.INFO_ESCALATED
    assert c.iterations == gp.last_result.iterations
    assert np.array_equal(c.mean.cpu().numpy(), mean.cpu().numpy())
    assert np.array_equal(c.var.cpu().numpy(), var.cpu().numpy())
    K = O.gram(spec, X)
    w_ref, _, _ = O.laplace_weight(K, y, O.PROBIT, lp, tol=1e-5)
    p_ref, _ = O.laplace_precision(K, y, O.PROBIT, lp, w_ref)
    m_ref, _ = O.predict(spec, X, Xt, w_ref, p_ref)
    assert rel_err(c.mean.cpu().numpy(), m_ref) < 1e-5
    for t in (1, 2):
        assert out[t].status == 0 and np.all(np.isfinite(out[t].mean.cpu().numpy()))


@pytest.mark.parametrize("count,capacity", [(4, 3), (40, 32)])
def test_serving_async_launch_finish_matches_step(count, capacity):
    """The asynchronous launch / finish pair gives bit-identical results to
    the synchronous step (items are reported one step later).  In the larger
    case every fifth request has N > 1024 and runs in the second lane, with
    its own step cadence."""
    from probit_amd.serving import ServingEngine
    shapes = [(150, 2, "probit", "EQ"), (300, 1, "gaussian", "Matern12"),
              (70, 3, "probit", "Matern12"), (260, 2, "gaussian", "EQ")]
    rng = np.random.default_rng(80)
    while len(shapes) < count:
        k = len(shapes)
        n = int(rng.integers(1100, 1400)) if k % 5 == 0 else int(rng.integers(40, 400))
        shapes.append((n, int(rng.integers(1, 4)), ("probit", "gaussian")[k % 2],
                       ("EQ", "Matern12")[k % 3 % 2]))
    probs = []
    for i, (N, D, lik, kind) in enumerate(shapes):
        X, y, theta, lp, spec = small_problem(60 + i, N, D, lik, kind)
        Xt = np.random.default_rng(70 + i).normal(size=(90, D))
        probs.append((X, y, theta, lp, lik, kind, Xt))

    def run(async_mode):
        eng = ServingEngine(precision="fp32", capacity=capacity, hbm_budget_bytes=1 << 33)
        out, queue = {}, list(range(count))
        while queue or eng.in_flight:
            launched = eng.launch() if async_mode else False
            while queue:
                X, y, theta, lp, lik, kind, Xt = probs[queue[0]]
                if not eng.submit(queue[0], X, y, kind, theta[0], theta[1], lik, lp[0],
                                  lp[1] if lik == "probit" else None, Xt):
                    break
                queue.pop(0)
            done = eng.finish() if async_mode and launched else ([] if async_mode else eng.step())
            for c in done:
                out[c.tag] = c
        eng.close()
        return out

    a, b = run(False), run(True)
    assert sorted(a) == sorted(b) == list(range(count))
    for t in range(count):
        assert a[t].iterations == b[t].iterations
        for k in ("mean", "var", "weight", "precision"):
            assert np.array_equal(getattr(a[t], k).cpu().numpy(), getattr(b[t], k).cpu().numpy())


@pytest.mark.parametrize("cls,bounds,noise", [("LaplaceGP", "single", 2.5),
                                              ("LaplaceGP", "double", 0.3),
                                              ("VBGP", "single", 2.5),
                                              ("VBGP", "double", 0.8)])
def test_closed_form_bounds_likelihood_end_to_end(cls, bounds, noise):
    """LaplaceGP / VBGP built with the closed-form derivative pair
    grad_/hessian_log_probit_likelihood (probit/utilities.py:151-192, as
    examples/classification.py:406-407 would pass them; functools.partial(...,
    single_precision=False) for BOUNDS["double"]): fit, predict and objective
    against the oracle, whose Newton / fixed point uses the same BOUNDS
    derivatives; the gradient is refused with NotImplementedError.  The
    reference's single bounds jump at |z| = 1.3 (its tail series is poor
    there), so its Newton cycles unless every |z| stays below 1.3 (sigma =
    2.5 here); the double bounds reach the tail regimes (|z| up to ~3.4 at
    sigma = 0.3) and converge."""
    import functools

    import probit_amd as P
    X, y, theta, lp, spec = small_problem(33, 160, 2, "probit", "EQ", noise=noise)
    g, h = P.grad_log_probit_likelihood, P.hessian_log_probit_likelihood
    kind = O.PROBIT_CLOSED
    if bounds == "double":
        g = functools.partial(g, single_precision=False)
        h = functools.partial(h, single_precision=False)
        kind = O.PROBIT_CLOSED_DOUBLE
    gp = getattr(P, cls)(data=(X, y), prior=make_prior("EQ"),
                         log_likelihood=P.log_probit_likelihood, grad_log_likelihood=g,
                         hessian_log_likelihood=h)
    params = (theta, lp)
    w, p = gp.approximate_posterior(params)
    K = O.gram(spec, X)
    if cls == "LaplaceGP":
        w_ref, it_ref, _ = O.laplace_weight(K, y, kind, lp, tol=1e-5)
        p_ref, _ = O.laplace_precision(K, y, kind, lp, w_ref)
        obj_ref = O.objective_LA(spec, X, y, kind, lp, w_ref)
    else:
        w_ref, it_ref, _ = O.vb_weight(K, y, kind, lp, tol=1e-5)
        p_ref = O.vb_precision(lp, 160)
        obj_ref = O.objective_VB(spec, X, y, kind, lp, w_ref)
    assert it_ref < 100 and gp.last_result.status == 0
    assert gp.last_result.iterations == it_ref
    assert rel_err(w.cpu().numpy(), w_ref) < 1e-9
    assert rel_err(p.cpu().numpy(), p_ref) < 1e-9
    Xt = np.random.default_rng(32).normal(size=(90, 2))
    mean, var = gp.predict(Xt, params, w, p)
    m_ref, v_ref = O.predict(spec, X, Xt, w_ref, p_ref)
    assert rel_err(mean.cpu().numpy(), m_ref) < 1e-9
    assert rel_err(var.cpu().numpy(), v_ref) < 1e-9
    assert abs(float(gp.objective()(params)) - obj_ref) <= 1e-9 * abs(obj_ref)
    with pytest.raises(NotImplementedError, match="closed-form"):
        gp.value_and_grad()(params)
