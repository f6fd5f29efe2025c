"""The C-ABI entry points that act on a resident factor (predict, covariance,
the two gradients) take the parameters they were called with, as the
reference's predict(X_test, parameters, ...) and value_and_grad()(parameters)
do (probit/approximators.py:154-180, :132-134): parameters other than the
resident fit's, or a workspace that holds another kind of state, are refused
with a negative status and a message, never answered from the resident
state.  Called through the C ABI."""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from tests.helpers import make_prior, small_problem

pytestmark = pytest.mark.gpu


def _fitted(cls="LaplaceGP", lik="probit"):
    import probit_amd as P
    X, y, theta, lp, _ = small_problem(60, 130, 2, lik, "EQ", noise=0.6)
    ll = P.log_gaussian_likelihood if lik == "gaussian" else P.log_probit_likelihood
    gp = getattr(P, cls)(data=(X, y), prior=make_prior("EQ"), log_likelihood=ll)
    params = (theta, lp)
    w, p = gp.approximate_posterior(params)
    return gp, params, w, p


def _predict(gp, ks, M=40):
    import torch

    from probit_amd import _lib
    lib = _lib.lib()
    Xt = torch.as_tensor(np.random.default_rng(61).normal(size=(M, 2)), device="cuda")
    mean = torch.empty(M, dtype=torch.float64, device="cuda")
    var = torch.empty_like(mean)
    need = lib.pgp_predict_workspace_size(gp.N, M, gp._prec)
    pws = _lib.workspace(need, "cuda")
    ws = gp._workspace()
    rc = lib.pgp_predict(ks, Xt.data_ptr(), M, mean.data_ptr(), var.data_ptr(), ws.data_ptr(),
                         ws.numel(), pws.data_ptr(), pws.numel(), gp._stream())
    cov = torch.empty((M, M), dtype=torch.float64, device="cuda")
    rc2 = lib.pgp_predict_covariance(ks, Xt.data_ptr(), M, cov.data_ptr(), M, ws.data_ptr(),
                                     ws.numel(), pws.data_ptr(), pws.numel(), gp._stream())
    return rc, rc2, mean


def _grad(gp, fn, ks, ls):
    from probit_amd import _lib
    lib = _lib.lib()
    out = np.zeros(2 + 2 * gp.D + _lib.MAX_CUT)
    ws, gws = gp._workspace(), gp._grad_workspace()
    return getattr(lib, fn)(ks, ls, out.ctypes.data, ws.data_ptr(), ws.numel(), gws.data_ptr(),
                            gws.numel(), gp._stream()), out


def _err():
    from probit_amd import _lib
    return _lib.load().pgp_last_error().decode()


def test_predict_refuses_other_kernel_parameters():
    gp, params, w, p = _fitted()
    ks, ls = gp._structs(params)
    rc, rc2, mean = _predict(gp, ks)
    assert rc == 0 and rc2 == 0
    for field, value in (("variance", 2.0), ("kind", 1)):
        other = gp._structs(params)[0]
        setattr(other, field, value)
        rc, rc2, _ = _predict(gp, other)
        assert rc == -1 and rc2 == -1, (field, rc, rc2)
        assert "differ from the resident fit" in _err()
    other = gp._structs(params)[0]
    other.lengthscale[1] = 0.5
    assert _predict(gp, other)[0] == -1
    other = gp._structs(params)[0]
    other.dim = 3
    assert _predict(gp, other)[0] == -1 and "dim" in _err()
    # the resident state is untouched: the right parameters still predict the same
    rc, _, mean2 = _predict(gp, ks)
    assert rc == 0 and bool((mean2 == mean).all())


def test_laplace_gradient_refuses_other_parameters_and_states():
    gp, params, w, p = _fitted()
    ks, ls = gp._structs(params)
    rc, g = _grad(gp, "pgp_laplace_grad", ks, ls)
    assert rc == 0 and np.all(np.isfinite(g))
    other = gp._structs(params)[1]
    other.noise_std = 0.61
    assert _grad(gp, "pgp_laplace_grad", ks, other)[0] == -2
    assert "likelihood parameters differ" in _err()
    other = gp._structs(params)[1]
    other.cutpoints[1] += 1e-3
    assert _grad(gp, "pgp_laplace_grad", ks, other)[0] == -2
    other_k = gp._structs(params)[0]
    other_k.lengthscale[0] *= 1.01
    assert _grad(gp, "pgp_laplace_grad", other_k, ls)[0] == -1
    # a Laplace state is not a VB fit
    assert _grad(gp, "pgp_vb_grad", ks, ls)[0] == -99
    assert "holds a Laplace state" in _err()


def test_vb_states_are_checked():
    gp, params, w, p = _fitted("VBGP")
    ks, ls = gp._structs(params)
    assert _grad(gp, "pgp_vb_grad", ks, ls)[0] == 0
    assert _grad(gp, "pgp_laplace_grad", ks, ls)[0] == -99
    assert "holds a VB fit" in _err()
    assert _predict(gp, ks)[0] == 0  # a VB fit is a predict state
    # after the VB map the workspace holds sigma^2 I + K: no predict from it
    gp.construct()(params, w)
    rc, rc2, _ = _predict(gp, ks)
    assert rc == -99 and rc2 == -99
    assert "VB map" in _err()
    # the Python layer re-forms the state it needs
    m, v = gp.predict(np.zeros((3, 2)), params, w, p)
    assert bool(torch_isfinite(m).all())


def torch_isfinite(t):
    import torch
    return torch.isfinite(t)
