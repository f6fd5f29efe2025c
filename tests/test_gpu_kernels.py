"""Kernel-level parity of the HIP library against the FP64 oracle, called
through the C ABI (probit_amd._lib)."""

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import gp_oracle as O
from tests.helpers import oracle_spec, rel_err

pytestmark = pytest.mark.gpu


def _setup():
    import torch

    from probit_amd import _lib
    return torch, _lib, _lib.lib()


def _kstruct(_lib, kind, var, ls, D, period=None):
    k = _lib.KernelT()
    k.kind = {"EQ": 0, "Matern12": 1}[kind]
    k.dim = D
    k.variance = var
    lsv = np.broadcast_to(np.asarray(ls, dtype=np.float64), (D,))
    for d in range(D):
        k.lengthscale[d] = lsv[d]
        if period is not None:
            k.period[d] = period
    return k


@pytest.mark.parametrize("kind", ["EQ", "Matern12"])
@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("ard", [False, True])
@pytest.mark.parametrize("dtype", [0, 1])
def test_gram_matches_oracle(kind, periodic, ard, dtype):
    torch, _lib, lib = _setup()
    rng = np.random.default_rng(0)
    N, M, D = 150, 70, 3
    X = rng.normal(size=(N, D))
    Z = rng.normal(size=(M, D))
    ls = np.array([0.7, 1.1, 1.9]) if ard else 0.9
    per = 0.8 if periodic else None
    spec = oracle_spec(kind, 1.4, ls, D, per)
    ks = _kstruct(_lib, kind, 1.4, ls, D, per)
    Xd = torch.as_tensor(X, device="cuda")
    Zd = torch.as_tensor(Z, device="cuda")
    tdt = torch.float64 if dtype == 0 else torch.float32
    tol = 2e-14 if dtype == 0 else 2e-6
    for (A, n, Bm, m) in [(Xd, N, Zd, M), (Xd, N, Xd, N)]:
        out = torch.empty((m, n), dtype=tdt, device="cuda")
        rc = lib.pgp_gram(ks, dtype, A.data_ptr(), n, Bm.data_ptr(), m, out.data_ptr(), n,
                          _lib.stream_handle())
        _lib.check(rc, "pgp_gram")
        torch.cuda.synchronize()
        got = out.t().double().cpu().numpy()
        ref = O.gram(spec, A.cpu().numpy(), Bm.cpu().numpy())
        assert np.max(np.abs(got - ref)) <= tol * 1.4, (kind, periodic, ard, dtype)


@pytest.mark.parametrize("lik", ["gaussian", "probit", "probit_closed"])
def test_likelihood_eval_matches_oracle(lik):
    torch, _lib, lib = _setup()
    rng = np.random.default_rng(1)
    N = 1000
    f = rng.normal(scale=3.0, size=N)
    ls = _lib.LikelihoodT()
    if lik == "gaussian":
        y = rng.normal(size=N)
        ls.kind = 0
        ls.noise_std = 0.4
        params = (0.4,)
    else:
        cut = np.array([-np.inf, -0.6, 0.1, 0.9, np.inf])
        y = rng.integers(0, 4, size=N).astype(np.float64)
        ls.kind = 1 if lik == "probit" else 2
        ls.J = 4
        ls.noise_std = 0.7
        ls.single_precision_bounds = 1
        for c in range(5):
            ls.cutpoints[c] = cut[c]
        params = (0.7, cut)
    fd = torch.as_tensor(f, device="cuda")
    yd = torch.as_tensor(y, device="cuda")
    ll, d1, d2 = (torch.empty_like(fd) for _ in range(3))
    rc = lib.pgp_likelihood_eval(ls, fd.data_ptr(), yd.data_ptr(), N, ll.data_ptr(),
                                 d1.data_ptr(), d2.data_ptr(), _lib.stream_handle())
    _lib.check(rc, "pgp_likelihood_eval")
    torch.cuda.synchronize()
    kind = {"gaussian": O.GAUSSIAN, "probit": O.PROBIT, "probit_closed": O.PROBIT_CLOSED}[lik]
    ref = O.lik_derivs(kind, f, y, params)
    for name, got in (("ll", ll), ("d1", d1), ("d2", d2)):
        g = got.cpu().numpy()
        r = ref[name]
        ok = np.isclose(g, r, rtol=1e-11, atol=1e-12) | (np.isnan(g) & np.isnan(r))
        assert ok.all(), (lik, name, np.max(np.abs(g - r)))


@pytest.mark.parametrize("single", [1, 0], ids=["single_bounds", "double_bounds"])
def test_closed_form_probit_every_regime_matches_oracle(single):
    """The BOUNDS derivatives at latent values that put each class interval's
    z-scores in every regime of _safe_Z (exact, two-sided and one-sided tails,
    linear; upper and lower; narrow, medium and wide classes, both ends
    unbounded for the extreme classes), against the oracle."""
    torch, _lib, lib = _setup()
    cut = np.array([-np.inf, -1.0, -0.8, 0.0, 2.0, np.inf])
    sigma = 0.5
    grid = np.linspace(-6.0, 8.0, 2801)
    f = np.tile(grid, 5)
    y = np.repeat(np.arange(5, dtype=np.float64), grid.size)
    N = f.size
    ls = _lib.LikelihoodT()
    ls.kind = 2
    ls.J = 5
    ls.noise_std = sigma
    ls.single_precision_bounds = single
    for c in range(6):
        ls.cutpoints[c] = cut[c]
    fd = torch.as_tensor(f, device="cuda")
    yd = torch.as_tensor(y, device="cuda")
    d1, d2 = torch.empty_like(fd), torch.empty_like(fd)
    rc = lib.pgp_likelihood_eval(ls, fd.data_ptr(), yd.data_ptr(), N, None, d1.data_ptr(),
                                 d2.data_ptr(), _lib.stream_handle())
    _lib.check(rc, "pgp_likelihood_eval")
    torch.cuda.synchronize()
    lp = (sigma, cut)
    r1 = O.grad_log_probit_likelihood(f, y, lp, single_precision=bool(single))
    r2 = O.hessian_log_probit_likelihood(f, y, lp, single_precision=bool(single))
    for name, got, r in (("d1", d1, r1), ("d2", d2, r2)):
        g = got.cpu().numpy()
        ok = np.isclose(g, r, rtol=1e-11, atol=1e-12) | (np.isnan(g) & np.isnan(r))
        bad = np.flatnonzero(~ok)
        assert ok.all(), (name, single, f[bad[:5]], y[bad[:5]], g[bad[:5]], r[bad[:5]])


def test_ordinal_metrics_match_oracle():
    """calculate_metrics (examples/classification.py:325-341) on the GPU: the
    argmax, the M x M P[:, y_test] matrix and the five numbers, against the
    oracle; a tie takes the first class, as argmax does."""
    torch, _lib, lib = _setup()
    import probit_amd as P
    rng = np.random.default_rng(9)
    M, J = 700, 4
    Pm = rng.dirichlet(np.ones(J), size=M)
    Pm[5] = [0.4, 0.4, 0.1, 0.1]  # tie: class 0
    y = rng.integers(0, J, size=M)
    zo, mae, lpp, plik = P.calculate_metrics(y, torch.as_tensor(Pm, device="cuda"),
                                             verbose=False)
    r_zo, r_mae, r_lpp, r_plik, r_pred = O.calculate_metrics(y, Pm)
    assert np.array_equal(plik.cpu().numpy(), r_plik)
    assert zo == r_zo and mae == r_mae
    assert abs(lpp - r_lpp) <= 1e-12 * abs(r_lpp)
    yp = torch.empty(M, dtype=torch.int32, device="cuda")
    met = torch.empty(5, dtype=torch.float64, device="cuda")
    ws = torch.empty(3 * M, dtype=torch.float64, device="cuda")
    Pd = torch.as_tensor(Pm, device="cuda")
    yd = torch.as_tensor(y.astype(np.float64), device="cuda")
    _lib.check(lib.pgp_ordinal_metrics(Pd.data_ptr(), J, yd.data_ptr(), M, yp.data_ptr(), None,
                                       met.data_ptr(), ws.data_ptr(), _lib.stream_handle()),
               "pgp_ordinal_metrics")
    assert np.array_equal(yp.cpu().numpy(), r_pred)
    assert yp.cpu().numpy()[5] == 0
    m = met.cpu().numpy()
    assert m[0] == np.sum(r_pred != y) and m[1] == np.sum(r_pred == y)


def _spd(rng, N):
    X = rng.normal(size=(N, 2))
    K = O.gram(O.KernelSpec.make(O.EQ, 1.0, 0.7, 2), X)
    s = np.sqrt(rng.uniform(0.5, 5.0, size=N))
    return np.eye(N) + s[:, None] * K * s[None, :]


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-5), (2, 3e-2)])
@pytest.mark.parametrize("N", [64, 256, 448])
def test_potrf_backward_error(prec, tol, N):
    torch, _lib, lib = _setup()
    rng = np.random.default_rng(N)
    A = _spd(rng, N)
    tdt = torch.float64 if prec == 0 else torch.float32
    Ad = torch.as_tensor(A, dtype=tdt, device="cuda").t().contiguous()  # column-major
    logd = torch.empty(N, dtype=torch.float64, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    dinv = torch.empty(N * 64, dtype=torch.float64, device="cuda")
    rc = lib.pgp_potrf(prec, Ad.data_ptr(), N, N, dinv.data_ptr(), logd.data_ptr(),
                       info.data_ptr(), _lib.stream_handle())
    _lib.check(rc, "pgp_potrf")
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    L = np.tril(Ad.t().double().cpu().numpy())
    Lref = np.linalg.cholesky(A)
    back = np.linalg.norm(L @ L.T - A) / np.linalg.norm(A)
    assert back < tol, (prec, N, back)
    if prec == 0:
        assert rel_err(L, Lref) < 1e-12
        assert rel_err(logd.cpu().numpy(), np.log(np.diag(Lref))) < 1e-12
    Dv = dinv.cpu().numpy().reshape(N // 64, 64, 64)  # block k, column-major
    for k in range(N // 64):
        blkL = L[64 * k:64 * k + 64, 64 * k:64 * k + 64]
        inv = Dv[k].T
        assert np.max(np.abs(inv @ blkL - np.eye(64))) < (1e-12 if prec == 0 else 1e-4)


def test_potrf_reports_not_pd():
    torch, _lib, lib = _setup()
    N = 128
    A = np.eye(N)
    A[70, 70] = -1.0
    Ad = torch.as_tensor(A, device="cuda").t().contiguous()
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    dinv = torch.empty(N * 64, dtype=torch.float64, device="cuda")
    rc = lib.pgp_potrf(0, Ad.data_ptr(), N, N, dinv.data_ptr(), None, info.data_ptr(),
                       _lib.stream_handle())
    _lib.check(rc, "pgp_potrf")
    torch.cuda.synchronize()
    assert int(info.item()) & _lib.INFO_NOT_PD


def test_potrf_rejects_bad_sizes():
    torch, _lib, lib = _setup()
    A = torch.zeros((100, 100), dtype=torch.float64, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    rc = lib.pgp_potrf(0, A.data_ptr(), 100, 100, A.data_ptr(), None, info.data_ptr(),
                       _lib.stream_handle())
    assert rc < 0


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("trans", [0, 1])
def test_trsv_matches_scipy(prec, trans):
    torch, _lib, lib = _setup()
    rng = np.random.default_rng(7)
    N = 320
    L = np.linalg.cholesky(_spd(rng, N))
    b = rng.normal(size=N)
    tdt = torch.float64 if prec == 0 else torch.float32
    Ld = torch.as_tensor(L, dtype=tdt, device="cuda").t().contiguous()
    Lh = Ld.t().double().cpu().numpy()
    bd = torch.as_tensor(b, device="cuda")
    x = torch.empty_like(bd)
    dinv = torch.empty(N * 64, dtype=torch.float64, device="cuda")
    rc = lib.pgp_trsv(prec, Ld.data_ptr(), N, N, dinv.data_ptr(), bd.data_ptr(), x.data_ptr(),
                      trans, _lib.stream_handle())
    _lib.check(rc, "pgp_trsv")
    torch.cuda.synchronize()
    ref = sla.solve_triangular(Lh, b, lower=True, trans=trans)
    assert rel_err(x.cpu().numpy(), ref) < (1e-12 if prec == 0 else 1e-12)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("N,M", [(256, 192), (640, 1088)])
def test_trsm_matches_scipy(prec, N, M):
    """Narrow right-hand sides take the per-block-row schedule, wide ones (>= 16
    column tiles) the super-step schedule; N = 640 ends in a partial
    super-step."""
    torch, _lib, lib = _setup()
    rng = np.random.default_rng(8)
    L = np.linalg.cholesky(_spd(rng, N))
    S = rng.normal(size=(N, M))
    tdt = torch.float64 if prec == 0 else torch.float32
    Ld = torch.as_tensor(L, dtype=tdt, device="cuda").t().contiguous()
    Sd = torch.as_tensor(S, dtype=tdt, device="cuda").t().contiguous()
    Lh = Ld.t().double().cpu().numpy()
    Sh = Sd.t().double().cpu().numpy()
    dinv = torch.empty(N * 64, dtype=torch.float64, device="cuda")
    rc = lib.pgp_trsm(prec, Ld.data_ptr(), N, N, dinv.data_ptr(), Sd.data_ptr(), M, N,
                      _lib.stream_handle())
    _lib.check(rc, "pgp_trsm")
    torch.cuda.synchronize()
    ref = sla.solve_triangular(Lh, Sh, lower=True)
    assert rel_err(Sd.t().double().cpu().numpy(), ref) < (1e-12 if prec == 0 else 1e-5)


def test_probit_predictive_matches_oracle():
    torch, _lib, lib = _setup()
    from probit_amd.utilities import probit_predictive_distributions
    rng = np.random.default_rng(9)
    mean = rng.normal(size=300)
    var = rng.uniform(0.01, 2.0, size=300)
    cut = np.array([-np.inf, -0.5, 0.2, np.inf])
    got = probit_predictive_distributions((0.6, cut), mean, var).cpu().numpy()
    ref = O.probit_predictive_distributions((0.6, cut), mean, var)
    assert np.max(np.abs(got - ref)) < 1e-14
    assert np.allclose(got.sum(axis=1), 1.0, atol=1e-14)


@pytest.mark.parametrize("N", [1600, 700])
def test_bf16_update_factor_log_determinant(N):
    """The BF16-update factor of B = I + W^1/2 K W^1/2 (pgp_laplace_set_state in
    the BF16 mode: super-steps of 8 block columns whose outer updates read the
    panel's BF16 copy; N = 700 has a single outer update) gives log|B| within
    BF16-update accuracy of the oracle's FP64 value, and is not the FP32-update
    factor's: a wrong operand in the BF16 path shows as a non-positive pivot
    or an O(1) error."""
    torch, _lib, lib = _setup()
    from probit_amd.utilities import likelihood_struct
    rng = np.random.default_rng(N)
    D = 3
    X = rng.normal(size=(N, D))
    cut = np.array([-np.inf, -0.3, 0.4, np.inf])
    y = np.searchsorted(cut[1:-1], np.sin(X.sum(1)) + 0.5 * rng.normal(size=N)).astype(np.float64)
    a = 0.05 * rng.normal(size=N)  # a weight away from zero: W varies over the points
    ks = _kstruct(_lib, "EQ", 1.2, 0.8, D)
    ls = likelihood_struct(_lib.LIK_PROBIT, (0.5, cut))
    Xd = torch.as_tensor(X, device="cuda")
    yd = torch.as_tensor(y, device="cuda")
    ad = torch.as_tensor(a, device="cuda")
    out = {}
    for prec in (0, 1, 2):
        W = torch.empty(N, dtype=torch.float64, device="cuda")
        prob = _lib.ProblemT(N, D, 0, Xd.data_ptr(), yd.data_ptr(), None, W.data_ptr(), None)
        ws = _lib.workspace(lib.pgp_workspace_size(N, D, prec), "cuda")
        res = _lib.FitResultT()
        rc = lib.pgp_laplace_set_state(prob, ks, ls, prec, ad.data_ptr(), None, res,
                                       ws.data_ptr(), ws.numel(), _lib.stream_handle())
        _lib.check(rc, "pgp_laplace_set_state")
        torch.cuda.synchronize()
        assert res.status == 0, (prec, res.status)
        out[prec] = (res.log_det, W.cpu().numpy())
    # the oracle's log|B| at the precision the library formed
    K = O.gram(O.KernelSpec.make(O.EQ, 1.2, 0.8, D), X)
    s = np.sqrt(np.maximum(out[0][1], 0.0))
    B = np.eye(N) + s[:, None] * K * s[None, :]
    ld = 2.0 * np.sum(np.log(np.diag(np.linalg.cholesky(B))))
    print(f"N {N}: log|B| {ld:.6f}; relative error fp64 {abs(out[0][0] - ld) / abs(ld):.2e} "
          f"fp32 {abs(out[1][0] - ld) / abs(ld):.2e} bf16 {abs(out[2][0] - ld) / abs(ld):.2e}",
          flush=True)
    assert abs(out[0][0] - ld) <= 1e-10 * abs(ld)
    # measured: FP32 4e-6 / 1e-6, BF16 8e-5 / 2e-4 (N = 1600 / 700)
    assert abs(out[1][0] - ld) <= 2e-5 * abs(ld)
    assert abs(out[2][0] - ld) <= 2e-3 * abs(ld), (out[2][0], ld)
    assert out[2][0] != out[1][0]
