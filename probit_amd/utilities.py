"""Likelihoods and helpers with the reference's names and signatures
(probit/utilities.py).

The likelihood functions are the identity tokens the approximators dispatch
on: passing ``log_gaussian_likelihood`` or ``log_probit_likelihood`` (and
optionally the closed-form ``grad_/hessian_log_probit_likelihood`` pair)
selects the matching fused HIP kernel inside the Newton step.  Called
directly, they evaluate elementwise on the GPU through the same kernel
(``pgp_likelihood_eval``), with the reference's argument convention
``f(f, y, likelihood_parameters)``.
"""

from __future__ import annotations

import math
import warnings

import numpy as np
import torch

from . import _lib

__all__ = [
    "BOUNDS",
    "h",
    "log_gaussian_likelihood",
    "log_probit_likelihood",
    "grad_log_probit_likelihood",
    "hessian_log_probit_likelihood",
    "probit_predictive_distributions",
    "calculate_metrics",
    "check_data",
    "check_cutpoints",
    "ndtr",
    "norm_cdf",
    "norm_pdf",
    "norm_z_pdf",
    "norm_logpdf",
    "norm_z_logpdf",
    "probit",
    "probit_likelihood",
    "CutpointValueError",
    "InvalidKernel",
]

BOUNDS = {"single": [1.3, 1.8, 2.3], "double": [2.3, 3.6, 4.8]}  # probit/utilities.py:15


def h(x):
    """Polynomial part of the normal log-survival tail series
    (probit/utilities.py:37-44): -x^-2 + 5/2 x^-4 - 37/3 x^-6.  A host
    helper; the device copy lives in the closed-form likelihood kernel."""
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return -1.0 * x**-2 + 5.0 / 2.0 * x**-4 - 37.0 / 3.0 * x**-6


# --------------------------------------------------------------------------
# elementwise normal / probit helpers (probit/utilities.py:18-34,47-53,64-70,
# 195-229): the reference's formulas on float64 tensors, on the device of
# their input.  The Newton / predict path does not call them; it evaluates
# the same quantities inside the fused HIP likelihood kernels.
# --------------------------------------------------------------------------
_OVER_SQRT_2PI = 1.0 / math.sqrt(2.0 * math.pi)


def _t(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(torch.float64)
    return torch.as_tensor(np.asarray(x, dtype=np.float64))


def ndtr(z):
    """Standard normal CDF as 0.5 (1 + erf(z / sqrt 2)) (probit/utilities.py:18-19)."""
    return 0.5 * (1.0 + torch.erf(_t(z) / math.sqrt(2.0)))


def norm_z_pdf(z):
    """probit/utilities.py:22-23."""
    z = _t(z)
    return _OVER_SQRT_2PI * torch.exp(-0.5 * z * z)


def norm_pdf(x, loc=0.0, scale=1.0):
    """probit/utilities.py:26-28."""
    return norm_z_pdf((_t(x) - loc) / scale) / scale


def norm_cdf(x):
    """ndtr with +-inf mapped to 1 / 0 (probit/utilities.py:31-34)."""
    x = _t(x)
    inf = torch.isinf(x)
    safe = torch.where(inf, torch.ones_like(x), x)
    return torch.where(inf, (x == math.inf).to(x.dtype), ndtr(safe))


def norm_z_logpdf(x):
    """probit/utilities.py:64-65."""
    x = _t(x)
    return math.log(_OVER_SQRT_2PI) - x * x / 2.0


def norm_logpdf(x, loc=0.0, scale=1.0):
    """probit/utilities.py:68-70."""
    return norm_z_logpdf((_t(x) - loc) / scale) - math.log(scale)


def probit(noise_std, cutpoints_y, cutpoints_yplus1, f):
    """Z = Phi((b_{y+1} - f) / s) - Phi((b_y - f) / s) with the infinite
    cutpoints guarded (probit/utilities.py:195-229)."""
    f = _t(f)
    b1 = _t(cutpoints_y).to(f.device)
    b2 = _t(cutpoints_yplus1).to(f.device)
    z1 = torch.where(b1 == -math.inf, torch.zeros_like(f), b1 - f)
    z2 = torch.where(b2 == math.inf, torch.zeros_like(f), b2 - f)
    c1 = torch.where(b1 == -math.inf, torch.zeros_like(f), norm_cdf(z1 / noise_std))
    c2 = torch.where(b2 == math.inf, torch.ones_like(f), norm_cdf(z2 / noise_std))
    return c2 - c1


def probit_likelihood(f, y, likelihood_parameters):
    """probit(noise_std, cutpoints[y], cutpoints[y + 1], f) (probit/utilities.py:47-53)."""
    f = _t(f)
    cut = _t(likelihood_parameters[1]).to(f.device)
    yi = _t(y).to(f.device).long()
    return probit(_scalar(likelihood_parameters[0]), cut[yi], cut[yi + 1], f)


# --------------------------------------------------------------------------
# likelihood parameter parsing
# --------------------------------------------------------------------------
def _scalar(v) -> float:
    if isinstance(v, torch.Tensor):
        return float(v.detach().reshape(()))
    return float(np.asarray(v, dtype=np.float64).reshape(()))


def likelihood_struct(kind: int, likelihood_parameters, single_precision: bool = True
                      ) -> _lib.LikelihoodT:
    """Pack likelihood parameters ((noise_std,) or (noise_std, cutpoints))."""
    lp = likelihood_parameters
    if not isinstance(lp, (tuple, list)):
        lp = (lp,)
    s = _lib.LikelihoodT()
    s.kind = kind
    s.noise_std = _scalar(lp[0])
    s.single_precision_bounds = 1 if single_precision else 0
    if kind != _lib.LIK_GAUSSIAN:
        if len(lp) < 2:
            raise ValueError("probit likelihood parameters are (noise_std, cutpoints)")
        cut = lp[1]
        cut = cut.detach().cpu().numpy() if isinstance(cut, torch.Tensor) else np.asarray(cut)
        cut = np.asarray(cut, dtype=np.float64).reshape(-1)
        J = cut.shape[0] - 1
        if J < 2 or J + 1 > _lib.MAX_CUT:
            raise ValueError(f"need 3..{_lib.MAX_CUT} cutpoints incl. +-inf (got {J + 1})")
        s.J = J
        for c in range(J + 1):
            s.cutpoints[c] = float(cut[c])
    return s


def _eval(kind, f, y, likelihood_parameters, which, single_precision=True):
    lib = _lib.lib()
    ft = torch.as_tensor(f, dtype=torch.float64)
    shape = ft.shape
    dev = ft.device if ft.is_cuda else torch.device("cuda")
    fd = ft.reshape(-1).to(dev).contiguous()
    yd = torch.as_tensor(y, dtype=torch.float64).reshape(-1).to(dev)
    yd = torch.broadcast_to(yd, fd.shape).contiguous()
    out = torch.empty_like(fd)
    ls = likelihood_struct(kind, likelihood_parameters, single_precision)
    ptrs = [None, None, None]
    ptrs[which] = out.data_ptr()
    rc = lib.pgp_likelihood_eval(ls, fd.data_ptr(), yd.data_ptr(), fd.numel(), *ptrs,
                                 _lib.stream_handle(dev))
    _lib.check(rc, "pgp_likelihood_eval")
    return out.reshape(shape)


def log_gaussian_likelihood(f, y, likelihood_parameters):
    """log N(y; f, noise_std^2) (probit/utilities.py:60-61)."""
    return _eval(_lib.LIK_GAUSSIAN, f, y, likelihood_parameters, 0)


def log_probit_likelihood(f, y, likelihood_parameters):
    """log(Phi(z2) - Phi(z1) + 1e-10) of the class interval (probit/utilities.py:56-57)."""
    return _eval(_lib.LIK_PROBIT, f, y, likelihood_parameters, 0)


def grad_log_probit_likelihood(f, y, likelihood_parameters, single_precision=True):
    """Closed-form gradient with BOUNDS tails (probit/utilities.py:151-169)."""
    return _eval(_lib.LIK_PROBIT_CLOSED, f, y, likelihood_parameters, 1, single_precision)


def hessian_log_probit_likelihood(f, y, likelihood_parameters, single_precision=True):
    """Closed-form second derivative with BOUNDS tails (probit/utilities.py:172-192)."""
    return _eval(_lib.LIK_PROBIT_CLOSED, f, y, likelihood_parameters, 2, single_precision)


def probit_predictive_distributions(likelihood_parameters, posterior_mean, posterior_variance):
    """(N_test, J) ordinal class probabilities (probit/utilities.py:232-249)."""
    lib = _lib.lib()
    mean = torch.as_tensor(posterior_mean, dtype=torch.float64)
    dev = mean.device if mean.is_cuda else torch.device("cuda")
    mean = mean.reshape(-1).to(dev).contiguous()
    var = torch.as_tensor(posterior_variance, dtype=torch.float64).reshape(-1).to(dev).contiguous()
    ls = likelihood_struct(_lib.LIK_PROBIT, likelihood_parameters)
    out = torch.empty((mean.numel(), ls.J), dtype=torch.float64, device=dev)
    rc = lib.pgp_probit_predictive(ls, mean.data_ptr(), var.data_ptr(), mean.numel(),
                                   out.data_ptr(), _lib.stream_handle(dev))
    _lib.check(rc, "pgp_probit_predictive")
    return out


def calculate_metrics(y_test, predictive_distributions, verbose=True):
    """Test metrics of the reference's classification example
    (examples/classification.py:325-341) on the GPU (pgp_ordinal_metrics):
    returns (mean_zero_one, mean_absolute_error, log_predictive_probability,
    predictive_likelihood) and, with ``verbose``, prints the reference's lines.

    As in the reference, ``predictive_likelihood`` is
    ``predictive_distributions[:, y_test]`` -- an (M, M) matrix, column j
    holding every test point's probability of class y_test[j] -- and the log
    predictive probability sums its log over all M * M entries; the numbers
    are the reference's, not the per-point log likelihood."""
    lib = _lib.lib()
    P = torch.as_tensor(predictive_distributions, dtype=torch.float64)
    dev = P.device if P.is_cuda else torch.device("cuda")
    P = P.to(dev).contiguous()
    if P.dim() != 2:
        raise ValueError("predictive_distributions must be (M, J)")
    M, J = P.shape
    y = torch.as_tensor(y_test).reshape(-1).to(device=dev, dtype=torch.float64).contiguous()
    if y.numel() != M:
        raise ValueError("y_test and predictive_distributions disagree on M")
    y_pred = torch.empty(M, dtype=torch.int32, device=dev)
    plik = torch.empty((M, M), dtype=torch.float64, device=dev)
    metrics = torch.empty(5, dtype=torch.float64, device=dev)
    ws = torch.empty(3 * M, dtype=torch.float64, device=dev)
    rc = lib.pgp_ordinal_metrics(P.data_ptr(), J, y.data_ptr(), M, y_pred.data_ptr(),
                                 plik.data_ptr(), metrics.data_ptr(), ws.data_ptr(),
                                 _lib.stream_handle(dev))
    _lib.check(rc, "pgp_ordinal_metrics")
    n_bad, n_good, mae, lpp, zero_one = metrics.cpu().tolist()
    if verbose:
        print(int(n_bad), "sum incorrect")
        print(int(n_good), "sum correct")
        print("mean_absolute_error={:.2f}".format(mae))
        print("log_pred_probability={:.2f}".format(lpp))
        print("mean_zero_one_error={:.2f}".format(zero_one))
    return zero_one, mae, lpp, plik


# --------------------------------------------------------------------------
# validators (probit/utilities.py:252-361), restated with working padding
# --------------------------------------------------------------------------
def check_data(data):
    """Class labels must be integers (probit/utilities.py:252-260)."""
    X_train, y_train = data
    dt = y_train.dtype
    is_int = (np.issubdtype(np.dtype(str(dt).replace("torch.", "")), np.integer)
              if not isinstance(dt, np.dtype) else np.issubdtype(dt, np.integer))
    if not is_int:
        raise TypeError(f"t must contain only integer values (got {dt})")
    return X_train, y_train.astype(int) if isinstance(y_train, np.ndarray) else y_train.long()


def check_cutpoints(cutpoints, J):
    """Pad / validate cutpoints to J+1 entries with -inf and +inf ends
    (probit/utilities.py:263-322)."""
    warnings.warn("Checking cutpoints...")
    c = cutpoints.detach().cpu().numpy() if isinstance(cutpoints, torch.Tensor) else cutpoints
    c = np.asarray(c, dtype=np.float64).reshape(-1)
    if c.shape[0] == J - 1:
        c = np.concatenate([[-np.inf], c, [np.inf]])
    elif c.shape[0] == J:
        if c[-1] != np.inf:
            if c[0] != -np.inf:
                raise ValueError(
                    "Either the largest cutpoint parameter b_J is not positive infinity, or the "
                    f"smallest cutpoint parameter b_0 is not negative infinity (got "
                    f"{[c[0], c[-1]]}, expected {[-np.inf, np.inf]})")
            c = np.concatenate([c, [np.inf]])
        else:
            c = np.concatenate([[-np.inf], c])
    elif c.shape[0] == J + 1:
        if c[0] != -np.inf:
            raise ValueError(
                f"The smallest cutpoint parameter b_0 must be negative infinity (got {c[0]})")
        if c[-1] != np.inf:
            raise ValueError(
                f"The largest cutpoint parameter b_J must be positive infinity (got {c[-1]})")
    else:
        raise ValueError(f"Could not recognise cutpoints shape. (shape was {c.shape})")
    if not all(c[i] <= c[i + 1] for i in range(J)):
        raise CutpointValueError(c)
    return c


class CutpointValueError(Exception):
    """An invalid cutpoint argument (probit/utilities.py:325-345)."""

    def __init__(self, cutpoint):
        super().__init__(
            f"The cutpoint list or array must be in ascending order,  {cutpoint} was given.")


class InvalidKernel(Exception):
    """The prior did not return a kernel (probit/utilities.py:348-361)."""

    def __init__(self, kernel):
        super().__init__(f"{kernel} is not an instance of probit_amd.kernels.Kernel")


SQRT2 = math.sqrt(2.0)
