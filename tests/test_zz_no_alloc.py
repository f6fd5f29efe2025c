"""pgp_laplace_fit_noalloc (the fit whose workspace is the library's static
device array) against pgp_laplace_fit with a caller workspace, through the C
ABI: bit for bit on two calls in a row, the second with a smaller problem
(the array keeps the first call's contents); sizes beyond the limit are
refused before any work."""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(seed, N, lik):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (N, 3))
    f = np.sin(2.5 * X.sum(1))
    if lik == "gaussian":
        return X, f + 0.2 * rng.normal(size=N), (0.2,)
    cut = np.array([-np.inf, -0.4, 0.3, np.inf])
    y = np.searchsorted(cut[1:-1], f + 0.5 * rng.normal(size=N)).astype(np.float64)
    return X, y, (0.5, cut)


def _fit(lib, _lib, torch, X, y, lp, lik, prec, noalloc):
    from probit_amd.utilities import likelihood_struct
    N, D = X.shape
    Xd = torch.as_tensor(X, device="cuda")
    yd = torch.as_tensor(y, device="cuda")
    w = torch.full((N,), math.nan, dtype=torch.float64, device="cuda")
    W = torch.full_like(w, math.nan)
    f = torch.full_like(w, math.nan)
    prob = _lib.ProblemT(N, D, 0, Xd.data_ptr(), yd.data_ptr(), w.data_ptr(), W.data_ptr(),
                         f.data_ptr())
    ks = _lib.KernelT()
    ks.kind, ks.dim, ks.variance = _lib.KERNEL_MATERN12, D, 1.3
    for d in range(D):
        ks.lengthscale[d] = 0.6 + 0.1 * d
    ls = likelihood_struct(_lib.LIK_GAUSSIAN if lik == "gaussian" else _lib.LIK_PROBIT, lp)
    opts = _lib.FitOptionsT(1e-5, 100, prec, -1, 0)
    res = _lib.FitResultT()
    stream = _lib.stream_handle()
    if noalloc:
        rc = lib.pgp_laplace_fit_noalloc(prob, ks, ls, opts, res, stream)
    else:
        ws = _lib.workspace(lib.pgp_workspace_size(N, D, prec), "cuda")
        rc = lib.pgp_laplace_fit(prob, ks, ls, opts, res, ws.data_ptr(), ws.numel(), stream)
    _lib.check(rc, "fit")
    torch.cuda.synchronize()
    return w, W, f, res


@pytest.mark.parametrize("prec", [0, 1, 2], ids=["fp64", "fp32", "bf16"])
def test_noalloc_fit_equals_workspace_fit_on_two_calls(prec):
    import torch

    from probit_amd import _lib
    lib = _lib.lib()
    for seed, N, lik in ((80, 1000, "probit"), (81, 333, "gaussian")):
        X, y, lp = _problem(seed, N, lik)
        print(f"noalloc prec {prec} N {N} {lik}", flush=True)
        a = _fit(lib, _lib, torch, X, y, lp, lik, prec, noalloc=True)
        b = _fit(lib, _lib, torch, X, y, lp, lik, prec, noalloc=False)
        for name, ta, tb in zip(("weight", "precision", "latent"), a[:3], b[:3]):
            assert torch.equal(ta, tb), (prec, N, name)
        ra, rb = a[3], b[3]
        assert ra.status == rb.status and ra.iterations == rb.iterations, (prec, N)
        for field in ("step_norm", "sum_log_lik", "half_f_dot_a", "log_det", "objective"):
            va, vb = getattr(ra, field), getattr(rb, field)
            assert va == vb or (math.isnan(va) and math.isnan(vb)), (prec, N, field)
        assert ra.status & ~_lib.INFO_ESCALATED == 0 and math.isfinite(ra.objective)


def test_noalloc_refuses_sizes_beyond_its_limit():
    import torch

    from probit_amd import _lib
    lib = _lib.lib()
    N = _lib.NOALLOC_MAX_NP + 1
    X, y, lp = _problem(82, N, "gaussian")
    with pytest.raises(RuntimeError, match="no-allocation limit"):
        _fit(lib, _lib, torch, X, y, lp, "gaussian", 0, noalloc=True)
    # the limit itself fits
    X, y, lp = _problem(83, _lib.NOALLOC_MAX_NP, "gaussian")
    w, _, _, res = _fit(lib, _lib, torch, X, y, lp, "gaussian", 1, noalloc=True)
    assert res.status == 0 and bool(torch.isfinite(w).all())
    assert lib.pgp_noalloc_scratch_bytes() >= lib.pgp_workspace_size(_lib.NOALLOC_MAX_NP, 3, 0)
