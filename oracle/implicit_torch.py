"""Torch float64 (CPU) restatement of the reference's evidence gradient.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference obtains d objective_LA / d parameters with jax.value_and_grad
through ``fixed_point_layer`` (probit/approximators.py:132-134, 248-263):
  * forward: the Newton fixed point z* of f_LA (probit/implicit/Laplace.py:4-9);
  * backward (probit/implicit/solvers.py:52-64): with z_bar the cotangent
    of z*, solve u = (d f_LA / d z)^T u + z_bar, then return
    (d f_LA / d params)^T u,
plus autodiff of objective_LA (probit/implicit/Laplace.py:12-30) itself.
This module restates exactly that structure with torch autograd (dense
Jacobian and dense solve for the adjoint fixed point), for the EQ kernel.
"""

from __future__ import annotations

import math

import numpy as np
import torch

T = torch.float64


def _gram(X, var, ls):
    d = X[:, None, :] - X[None, :, :]
    return var * torch.exp(-0.5 * (d * d).sum(-1) / ls**2)


def _loglik(lik, f, y, sig, cut):
    if lik == "gaussian":
        z = (f - y) / sig
        return -0.5 * math.log(2 * math.pi) - z * z / 2 - torch.log(sig)
    yi = torch.as_tensor(y, dtype=torch.long)
    lo, hi = cut[yi], cut[yi + 1]
    lo_inf, hi_inf = torch.isinf(lo), torch.isinf(hi)

    def ndtr(z):
        return 0.5 * (1 + torch.erf(z / math.sqrt(2.0)))

    # "safe where" guards as probit/utilities.py:220-228
    lo_s = torch.where(lo_inf, torch.zeros_like(lo), lo)
    hi_s = torch.where(hi_inf, torch.zeros_like(hi), hi)
    c1 = torch.where(lo_inf, torch.zeros_like(f), ndtr((lo_s - f) / sig))
    c2 = torch.where(hi_inf, torch.ones_like(f), ndtr((hi_s - f) / sig))
    return torch.log(c2 - c1 + 1e-10)


def _dlik(lik, f, y, sig, cut, order):
    """Elementwise derivatives in f (the vmapped grad / grad-of-grad of
    probit/approximators.py:92-104), keeping the graph for outer autodiff."""
    fx = f if f.requires_grad else f.clone().requires_grad_(True)
    ll = _loglik(lik, fx, y, sig, cut)
    d1, = torch.autograd.grad(ll.sum(), fx, create_graph=True)
    if order == 1:
        return d1, fx
    d2, = torch.autograd.grad(d1.sum(), fx, create_graph=True)
    return d2, fx


def implicit_gradient(X, y, lik, variance, lengthscale, lp, weight):
    Xt = torch.as_tensor(X, dtype=T)
    var = torch.tensor(float(variance), dtype=T, requires_grad=True)
    ls = torch.tensor(float(lengthscale), dtype=T, requires_grad=True)
    sig = torch.tensor(float(lp[0]), dtype=T, requires_grad=True)
    params = [var, ls, sig]
    cut = None
    if lik != "gaussian":
        cut = torch.tensor(np.asarray(lp[1], dtype=np.float64), dtype=T, requires_grad=True)
        params.append(cut)
    yt = torch.as_tensor(y, dtype=T)

    def f_LA(z, var, ls, sig, cut):
        f = _gram(Xt, var, ls) @ z
        d1, _ = _dlik(lik, f, yt, sig, cut, 1)
        return d1

    def objective_LA(z, var, ls, sig, cut):
        K = _gram(Xt, var, ls)
        f = K @ z
        W = -_dlik(lik, f, yt, sig, cut, 2)[0]
        Lc = torch.linalg.cholesky(_gram(Xt, var, ls) + torch.diag(1.0 / W))
        return (-_loglik(lik, f, yt, sig, cut).sum() + 0.5 * f @ z
                + torch.log(torch.diagonal(Lc)).sum() + 0.5 * torch.log(W).sum())

    z = torch.as_tensor(weight, dtype=T).clone().requires_grad_(True)
    obj = objective_LA(z, var, ls, sig, cut)
    explicit = torch.autograd.grad(obj, params + [z], allow_unused=True)
    zbar = explicit[-1]
    explicit = explicit[:-1]
    # adjoint fixed point u = (d f_LA / d z)^T u + z_bar  (solvers.py:59-63)
    zd = z.detach()
    Jz = torch.autograd.functional.jacobian(
        lambda zz: f_LA(zz, var.detach(), ls.detach(), sig.detach(),
                        None if cut is None else cut.detach()), zd)
    u = torch.linalg.solve(torch.eye(zd.shape[0], dtype=T) - Jz.T, zbar)
    # (d f_LA / d params)^T u  (solvers.py:54)
    fz = f_LA(zd, var, ls, sig, cut)
    implicit = torch.autograd.grad(fz, params, grad_outputs=u, allow_unused=True)
    tot = [(e if e is not None else 0) + (i if i is not None else 0)
           for e, i in zip(explicit, implicit)]
    out = {"variance": float(tot[0]), "lengthscale": float(tot[1]), "noise_std": float(tot[2])}
    if cut is not None:
        out["cutpoints"] = tot[3].detach().numpy()
    return out
