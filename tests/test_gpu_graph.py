"""The single-problem Laplace fit replays a captured HIP graph of its Newton
iteration once the iteration's launch structure recurs in a later fit on the
same workspace (probit/approximators.py:204-210 per fit; the iteration is
probit/implicit/Laplace.py:16-30).  Replays give the bits of the direct
launches: repeated fits on one workspace -- the same problem, then another
problem of the same size, with eager work of the program's own between the
fits -- are checked one by one against fits with PGP_FIT_NO_GRAPH, and the
library's counters show that the graph path ran."""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(seed, N, lik):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (N, 2))
    f = np.sin(2.5 * X.sum(1))
    if lik == "gaussian":
        return X, f + 0.2 * rng.normal(size=N), (0.2,)
    cut = np.array([-np.inf, -0.4, 0.3, np.inf])
    y = np.searchsorted(cut[1:-1], f + 0.5 * rng.normal(size=N)).astype(np.float64)
    return X, y, (0.5, cut)


def _stats(lib):
    v = [C.c_int64() for _ in range(3)]
    assert lib.pgp_fit_graph_stats(*[C.byref(x) for x in v]) == 0
    return [x.value for x in v]


def _fit(lib, _lib, torch, X, y, lp, lik, prec, ws, flags):
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
    ks.kind, ks.dim, ks.variance = _lib.KERNEL_EQ, D, 1.1
    for d in range(D):
        ks.lengthscale[d] = 0.5 + 0.2 * d
    ls = likelihood_struct(_lib.LIK_GAUSSIAN if lik == "gaussian" else _lib.LIK_PROBIT, lp)
    opts = _lib.FitOptionsT(1e-5, 100, prec, -1, flags)
    res = _lib.FitResultT()
    rc = lib.pgp_laplace_fit(prob, ks, ls, opts, res, ws.data_ptr(), ws.numel(),
                             _lib.stream_handle())
    _lib.check(rc, "fit")
    torch.cuda.synchronize()
    return (w, W, f), res


def _same(a, b, key):
    (ta, ra), (tb, rb) = a, b
    for name, x, y in zip(("weight", "precision", "latent"), ta, tb):
        assert __import__("torch").equal(x, y), (key, name)
    assert ra.status == rb.status and ra.iterations == rb.iterations, key
    for field in ("step_norm", "sum_log_lik", "half_f_dot_a", "log_det", "objective"):
        va, vb = getattr(ra, field), getattr(rb, field)
        assert va == vb or (math.isnan(va) and math.isnan(vb)), (key, field)


@pytest.mark.parametrize("prec", [0, 1, 2], ids=["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("lik,N", [("probit", 700), ("gaussian", 1300), ("probit", 100)])
def test_graph_replays_equal_direct_launches(prec, lik, N):
    import torch

    from probit_amd import _lib
    lib = _lib.lib()
    nbytes = lib.pgp_workspace_size(N, 2, prec)
    ws = _lib.workspace(nbytes, "cuda")
    other_ws = _lib.workspace(lib.pgp_workspace_size(300, 2, prec), "cuda")
    direct_ws = _lib.workspace(nbytes, "cuda")
    X, y, lp = _problem(90 + N, N, lik)
    X2, y2, lp2 = _problem(91 + N, N, lik)
    Xo, yo, lpo = _problem(92, 300, "probit")
    print(f"graph prec {prec} {lik} N {N}", flush=True)
    ref = _fit(lib, _lib, torch, X, y, lp, lik, prec, direct_ws, _lib.FIT_NO_GRAPH)
    ref2 = _fit(lib, _lib, torch, X2, y2, lp2, lik, prec, direct_ws, _lib.FIT_NO_GRAPH)
    s0 = _stats(lib)
    rng = torch.Generator(device="cuda").manual_seed(5)
    for rep in range(4):
        got = _fit(lib, _lib, torch, X, y, lp, lik, prec, ws, 0)
        _same(got, ref, (prec, lik, N, rep))
        # eager work of the program's own between the fits: other kernels,
        # another fit (fresh epochs, its own tickets) and a synchronisation
        a = torch.randn(256, 256, device="cuda", generator=rng)
        (a @ a.T).sum().item()
        _fit(lib, _lib, torch, Xo, yo, lpo, "probit", prec, other_ws, _lib.FIT_NO_GRAPH)
    # another problem of the same size on the same workspace replays the
    # graphs with new data
    for rep in range(2):
        _same(_fit(lib, _lib, torch, X2, y2, lp2, lik, prec, ws, 0), ref2, (prec, lik, N, "p2"))
    s1 = _stats(lib)
    replayed = s1[1] - s0[1]
    # the first fit on the workspace launches directly; from the second on,
    # every iteration's key has been seen and is replayed
    assert replayed >= 3 * ref[1].iterations, (s0, s1)
    assert ref[1].status & ~_lib.INFO_ESCALATED == 0


def test_no_graph_flag_launches_directly():
    import torch

    from probit_amd import _lib
    lib = _lib.lib()
    N = 400
    ws = _lib.workspace(lib.pgp_workspace_size(N, 2, 1), "cuda")
    X, y, lp = _problem(95, N, "probit")
    s0 = _stats(lib)
    for _ in range(3):
        _, res = _fit(lib, _lib, torch, X, y, lp, "probit", 1, ws, _lib.FIT_NO_GRAPH)
    s1 = _stats(lib)
    assert s1[0] == s0[0] and s1[1] == s0[1]
    assert s1[2] - s0[2] == 3 * res.iterations


def test_graph_cache_follows_the_schedule_table():
    """A schedule-table change gives the iteration a new launch structure: the
    graph captured under the old entry is not replayed for it (the cache is
    keyed by the entry), and the fit equals a direct fit under the new entry
    bit for bit."""
    import torch

    from probit_amd import _lib, autotune
    lib = _lib.lib()
    N, prec = 900, 1
    ws = _lib.workspace(lib.pgp_workspace_size(N, 2, prec), "cuda")
    direct_ws = _lib.workspace(lib.pgp_workspace_size(N, 2, prec), "cuda")
    X, y, lp = _problem(97, N, "probit")
    saved = autotune.get("fp32", 0)
    try:
        for _ in range(3):  # captured from the second fit on
            _fit(lib, _lib, torch, X, y, lp, "probit", prec, ws, 0)
        autotune.set_entry("fp32", 0, dict(saved, superstep=4 if saved["superstep"] != 4 else 6))
        ref = _fit(lib, _lib, torch, X, y, lp, "probit", prec, direct_ws, _lib.FIT_NO_GRAPH)
        for rep in range(3):
            _same(_fit(lib, _lib, torch, X, y, lp, "probit", prec, ws, 0), ref, ("new entry", rep))
    finally:
        autotune.set_entry("fp32", 0, saved)


def test_full_cache_evicts_the_least_recently_used_keys():
    """More keys than the cache holds (40 live workspaces, two or three keys
    each): a new workspace's iterations still get captured and replayed --
    the oldest keys make room -- and equal direct launches bit for bit; the
    first workspace, whose keys were evicted, fits correctly again."""
    import torch

    from probit_amd import _lib
    lib = _lib.lib()
    N, prec = 128, 1
    size = lib.pgp_workspace_size(N, 2, prec)
    X, y, lp = _problem(98, N, "probit")
    ref = _fit(lib, _lib, torch, X, y, lp, "probit", prec, _lib.workspace(size, "cuda"),
               _lib.FIT_NO_GRAPH)
    wss = [_lib.workspace(size, "cuda") for _ in range(40)]
    for ws in wss:
        for _ in range(2):
            _same(_fit(lib, _lib, torch, X, y, lp, "probit", prec, ws, 0), ref, "filling")
    ws = _lib.workspace(size, "cuda")
    s0 = _stats(lib)
    for rep in range(3):
        _same(_fit(lib, _lib, torch, X, y, lp, "probit", prec, ws, 0), ref, ("new ws", rep))
    s1 = _stats(lib)
    assert s1[1] - s0[1] >= 2 * ref[1].iterations, (s0, s1)
    for rep in range(2):
        _same(_fit(lib, _lib, torch, X, y, lp, "probit", prec, wss[0], 0), ref, ("first ws", rep))
