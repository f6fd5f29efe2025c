"""Covariance functions with mlkernels' compositional surface.

The reference builds its prior as, e.g.,
``signal_variance * EQ().stretch(lengthscale).periodic(0.5)``
(examples/regression.py:120-123, examples/classification.py:389-391) and
calls it as ``kernel(X)``, ``kernel(X, X_test)`` and ``kernel.elwise(X, X)``
(probit/approximators.py:171-174).  This module re-provides that surface for
the two families the GP path uses, EQ and Matern12, with scalar scaling,
scalar or per-dimension (ARD) ``stretch`` and ``periodic``.

Semantics (mlkernels): ``k.stretch(l)(x, y) = k(x / l, y / l)`` and
``k.periodic(p)(x, y) = k(phi(x), phi(y))`` with the feature map
``phi(x) = (sin(2 pi x / p), cos(2 pi x / p))`` per dimension, so
``EQ().stretch(l).periodic(p)`` has ``r^2 = sum_d 4 sin^2(pi dx_d / p) / l^2``.
A stretch applied on top of a periodic kernel rescales its period.

Every kernel reduces to ``variance * kappa(r^2)`` with per-dimension
lengthscales and (optional) periods, which is what the HIP Gram kernels
evaluate.  Parameter values may be Python floats, NumPy arrays or torch
tensors; tensors keep their autograd history, so gradients of the evidence
flow back to whatever parameterisation the caller used.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib

__all__ = ["Kernel", "EQ", "Matern12"]


def _as_tensor(v) -> torch.Tensor:
    if isinstance(v, torch.Tensor):
        return v.to(torch.float64)
    return torch.as_tensor(np.asarray(v, dtype=np.float64))


class Kernel:
    """variance * kappa(r^2) with per-dimension lengthscales / periods."""

    def __init__(self, kind: int, variance=1.0, lengthscale=1.0, period=None):
        self.kind = kind
        self.variance = variance
        self.lengthscale = lengthscale
        self.period = period

    # --- composition -----------------------------------------------------
    def _copy(self, **kw) -> "Kernel":
        d = dict(kind=self.kind, variance=self.variance, lengthscale=self.lengthscale,
                 period=self.period)
        d.update(kw)
        return Kernel(**d)

    def stretch(self, lengthscale) -> "Kernel":
        if self.period is None:
            return self._copy(lengthscale=_mul(self.lengthscale, lengthscale))
        return self._copy(period=_mul(self.period, lengthscale))

    def periodic(self, period) -> "Kernel":
        if self.period is not None:
            raise NotImplementedError("a periodic kernel cannot be made periodic again")
        return self._copy(period=period)

    def __mul__(self, other):
        if isinstance(other, Kernel):
            raise NotImplementedError("products of kernels are not supported")
        return self._copy(variance=_mul(self.variance, other))

    __rmul__ = __mul__

    def __add__(self, other):
        raise NotImplementedError("sums of kernels are not supported")

    __radd__ = __add__

    def __repr__(self):
        name = {_lib.KERNEL_EQ: "EQ", _lib.KERNEL_MATERN12: "Matern12"}[self.kind]
        s = f"{self.variance!r} * {name}().stretch({self.lengthscale!r})"
        if self.period is not None:
            s += f".periodic({self.period!r})"
        return s

    # --- numeric views -----------------------------------------------------
    def spec_tensors(self, D: int):
        """(variance (), lengthscale (D,), period (D,) or None) as float64 tensors."""
        var = _as_tensor(self.variance).reshape(())
        ls = _as_tensor(self.lengthscale)
        ls = ls.reshape(-1)
        if ls.numel() == 1:
            ls = ls.expand(D)
        elif ls.numel() != D:
            raise ValueError(f"lengthscale has {ls.numel()} entries for D={D}")
        per = None
        if self.period is not None:
            per = _as_tensor(self.period).reshape(-1)
            if per.numel() == 1:
                per = per.expand(D)
            elif per.numel() != D:
                raise ValueError(f"period has {per.numel()} entries for D={D}")
        return var, ls, per

    def to_struct(self, D: int) -> _lib.KernelT:
        if D < 1 or D > _lib.MAX_DIM:
            raise ValueError(f"input dimension must be in [1, {_lib.MAX_DIM}] (got {D})")
        var, ls, per = self.spec_tensors(D)
        k = _lib.KernelT()
        k.kind = self.kind
        k.dim = D
        k.variance = float(var.detach())
        lsv = ls.detach().cpu().numpy()
        for d in range(D):
            k.lengthscale[d] = float(lsv[d])
        if per is not None:
            pv = per.detach().cpu().numpy()
            for d in range(D):
                k.period[d] = float(pv[d])
        return k

    # --- evaluation (on the GPU) -------------------------------------------
    def __call__(self, X, Z=None) -> torch.Tensor:
        """Dense Gram matrix k(X, Z) (N x M, float64, on the GPU)."""
        lib = _lib.lib()
        Xd = _device_matrix(X)
        Zd = Xd if Z is None else _device_matrix(Z, Xd.device)
        N, D = Xd.shape
        M = Zd.shape[0]
        out = torch.empty((M, N), dtype=torch.float64, device=Xd.device)  # column-major N x M
        ks = self.to_struct(D)
        rc = lib.pgp_gram(ks, 0, Xd.data_ptr(), N, Zd.data_ptr(), M, out.data_ptr(), N,
                          _lib.stream_handle(Xd.device))
        _lib.check(rc, "pgp_gram")
        return out.t()

    def elwise(self, X, Z=None) -> torch.Tensor:
        """Diagonal k(x_i, x_i) for stationary kernels: the variance."""
        Xd = _device_matrix(X)
        var = float(_as_tensor(self.variance).detach())
        return torch.full((Xd.shape[0], 1), var, dtype=torch.float64, device=Xd.device)


def _mul(a, b):
    if isinstance(a, torch.Tensor) or isinstance(b, torch.Tensor):
        return _as_tensor(a) * _as_tensor(b)
    return np.asarray(a, dtype=np.float64) * np.asarray(b, dtype=np.float64) \
        if (np.ndim(a) or np.ndim(b)) else float(a) * float(b)


def _device_matrix(X, device=None) -> torch.Tensor:
    t = X if isinstance(X, torch.Tensor) else torch.as_tensor(np.asarray(X))
    t = t.detach().to(torch.float64)
    if t.dim() == 1:
        t = t[:, None]
    dev = device if device is not None else (t.device if t.is_cuda else torch.device("cuda"))
    return t.to(dev).contiguous()


def EQ() -> Kernel:
    """Exponentiated quadratic exp(-r^2 / 2) (mlkernels EQ)."""
    return Kernel(_lib.KERNEL_EQ)


def Matern12() -> Kernel:
    """Matern-1/2 exp(-r) (mlkernels Matern12)."""
    return Kernel(_lib.KERNEL_MATERN12)
