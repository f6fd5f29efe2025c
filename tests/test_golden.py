"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle) against the oracle (CPU) and
against the HIP path (GPU)."""

from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import gp_oracle as O
from tests.helpers import rel_err

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LAPLACE_CASES = ["regression_1d", "ordinal_matern", "ordinal_ard"]


def load(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def spec_of(g):
    per = g["period"] if g["period"].size else None
    return O.KernelSpec(int(g["kernel_kind"]), float(g["variance"]), g["lengthscale"], per)


def lp_of(g):
    if str(g["lik"]) == "gaussian":
        return (float(g["noise_std"]),)
    return (float(g["noise_std"]), g["cutpoints"])


def kind_of(g):
    return O.GAUSSIAN if str(g["lik"]) == "gaussian" else O.PROBIT


# ---------------------------------------------------------------- CPU: oracle
@pytest.mark.parametrize("name", LAPLACE_CASES)
def test_oracle_reproduces_laplace_fixture(name):
    g = load(name)
    spec, lp, kind = spec_of(g), lp_of(g), kind_of(g)
    K = O.gram(spec, g["X"])
    w, it, _ = O.laplace_weight(K, g["y"], kind, lp, tol=1e-5)
    assert it == int(g["iterations"])
    assert rel_err(w, g["weight"]) < 1e-10
    W, f = O.laplace_precision(K, g["y"], kind, lp, w)
    assert rel_err(W, g["precision"]) < 1e-10
    m, v = O.predict(spec, g["X"], g["Xt"], w, W)
    assert rel_err(m, g["mean"]) < 1e-10 and rel_err(v, g["var"]) < 1e-10
    obj = O.objective_LA(spec, g["X"], g["y"], kind, lp, w)
    assert abs(obj - float(g["objective"])) <= 1e-10 * abs(obj)


def test_oracle_reproduces_vb_fixture():
    g = load("vb_probit")
    spec, lp = spec_of(g), lp_of(g)
    K = O.gram(spec, g["X"])
    w, it, _ = O.vb_weight(K, g["y"], O.PROBIT, lp, tol=1e-5)
    assert it == int(g["iterations"]) and rel_err(w, g["weight"]) < 1e-10


def test_oracle_reproduces_likelihood_table():
    t = load("likelihoods")
    lp = (float(t["noise_std"]), t["cutpoints"])
    for tag, kind in (("exact", O.PROBIT), ("closed", O.PROBIT_CLOSED)):
        d = O.lik_derivs(kind, t["f"], t["y"], lp)
        for k in ("ll", "d1", "d2"):
            assert np.array_equal(np.isnan(d[k]), np.isnan(t[f"{tag}_{k}"]))
            ok = ~np.isnan(d[k])
            assert np.allclose(d[k][ok], t[f"{tag}_{k}"][ok], rtol=1e-13, atol=1e-300)


def test_fixture_detects_a_planted_oracle_mistake(monkeypatch):
    """A 1% error planted in the oracle's likelihood gradient moves the fixed
    point far outside the comparison tolerance: the fixtures pin behaviour."""
    g = load("ordinal_matern")
    spec, lp = spec_of(g), lp_of(g)
    K = O.gram(spec, g["X"])
    real = O.lik_derivs

    def planted(kind, f, y, lik, order=2):
        d = dict(real(kind, f, y, lik, order))
        d["d1"] = 1.01 * d["d1"]
        return d

    monkeypatch.setattr(O, "lik_derivs", planted)
    w, _, _ = O.laplace_weight(K, g["y"], O.PROBIT, lp, tol=1e-5)
    assert rel_err(w, g["weight"]) > 1e-4


# ---------------------------------------------------------------- GPU: HIP path
def _gp(cls, g, tol, precision_mode="fp64"):
    import probit_amd as P
    base = P.EQ if int(g["kernel_kind"]) == O.EQ else P.Matern12
    ls = g["lengthscale"]
    per = float(g["period"][0]) if g["period"].size else None

    def prior(theta):
        lsv, var = theta
        k = var * base().stretch(lsv)
        return k.periodic(per) if per is not None else k

    ll = P.log_gaussian_likelihood if str(g["lik"]) == "gaussian" else P.log_probit_likelihood
    gp = getattr(P, cls)(data=(g["X"], g["y"]), prior=prior, log_likelihood=ll, tolerance=tol,
                         precision_mode=precision_mode)
    theta = (ls if ls.size > 1 else float(ls[0]), float(g["variance"]))
    return gp, (theta, lp_of(g))


@pytest.mark.gpu
@pytest.mark.parametrize("name", LAPLACE_CASES)
def test_hip_laplace_matches_fixture(name):
    g = load(name)
    gp, params = _gp("LaplaceGP", g, 1e-5)
    w, p = gp.approximate_posterior(params)
    assert gp.last_result.iterations == int(g["iterations"])
    assert rel_err(w.cpu().numpy(), g["weight"]) < 1e-9
    assert rel_err(p.cpu().numpy(), g["precision"]) < 1e-9
    mean, var = gp.predict(g["Xt"], params, w, p)
    assert rel_err(mean.cpu().numpy(), g["mean"]) < 1e-9
    assert rel_err(var.cpu().numpy(), g["var"]) < 1e-9
    obj = float(gp.objective()(params))
    assert abs(obj - float(g["objective"])) <= 1e-9 * abs(obj)


@pytest.mark.gpu
@pytest.mark.parametrize("name", LAPLACE_CASES)
def test_hip_gradient_matches_fixture(name):
    g = load(name)
    gp, params = _gp("LaplaceGP", g, 1e-11)
    val, grads = gp.value_and_grad()(params)
    assert abs(val - float(g["objective_converged"])) <= 1e-9 * abs(val)
    (g_ls, g_var), g_lik = grads
    D = g["X"].shape[1]
    ref_ls = np.array([g[f"grad_lengthscale{d}"] for d in range(D)])
    got_ls = np.atleast_1d(np.asarray(g_ls, dtype=np.float64))
    if got_ls.size == 1 and D > 1:  # scalar lengthscale shared by all dimensions
        ref_ls = np.array([ref_ls.sum()])
    assert np.allclose(got_ls, ref_ls, rtol=1e-7, atol=1e-9)
    assert abs(g_var - float(g["grad_variance"])) <= 1e-7 * max(1.0, abs(float(g["grad_variance"])))
    assert abs(g_lik[0] - float(g["grad_noise_std"])) <= 1e-7 * max(1.0, abs(float(g["grad_noise_std"])))
    if "grad_cutpoints" in g:
        assert np.allclose(g_lik[1], g["grad_cutpoints"], rtol=1e-7, atol=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("name", LAPLACE_CASES)
def test_hip_fp32_mean_within_tolerance_of_fixture(name):
    g = load(name)
    gp, params = _gp("LaplaceGP", g, 1e-5, precision_mode="fp32")
    w, p = gp.approximate_posterior(params)
    mean, _ = gp.predict(g["Xt"], params, w, p)
    assert rel_err(mean.cpu().numpy(), g["mean"]) < 1e-5


@pytest.mark.gpu
def test_hip_vb_matches_fixture():
    g = load("vb_probit")
    gp, params = _gp("VBGP", g, 1e-5)
    w, p = gp.approximate_posterior(params)
    assert gp.last_result.iterations == int(g["iterations"])
    assert rel_err(w.cpu().numpy(), g["weight"]) < 1e-8
    assert np.allclose(p.cpu().numpy(), g["precision"], rtol=1e-15)
    obj = float(gp.objective()(params))
    assert abs(obj - float(g["objective"])) <= 1e-8 * abs(obj)


@pytest.mark.gpu
def test_hip_likelihoods_match_fixture():
    from probit_amd import _lib
    from probit_amd.utilities import _eval
    t = load("likelihoods")
    lp = (float(t["noise_std"]), t["cutpoints"])
    cases = [("exact", _lib.LIK_PROBIT, t["y"], lp), ("closed", _lib.LIK_PROBIT_CLOSED, t["y"], lp),
             ("gauss", _lib.LIK_GAUSSIAN, t["gauss_y"], (float(t["gauss_noise_std"]),))]
    for tag, kind, y, params in cases:
        for which, k in enumerate(("ll", "d1", "d2")):
            ref = t[f"{tag}_{k}"]
            got = _eval(kind, t["f"], y, params, which).cpu().numpy()
            ok = np.isclose(got, ref, rtol=1e-11, atol=1e-12) | (np.isnan(got) & np.isnan(ref))
            assert ok.all(), (tag, k, np.max(np.abs(got - ref)))
