"""The serving-stream configuration itself (N log-uniform on [64, 4096], D in
1..8, Gaussian or ordinal probit with J in 2..5, EQ / Matern12, 256 test
points; seed 4, as the benchmark serves it): a fixed slice of the stream --
its first 36 requests and its largest ones (N = 4096 with D = 8; N ~ 4000
with J = 5) -- served through ServingEngine's asynchronous launch / finish
in every precision mode, each completion checked against the oracle
(probit/approximators.py:204-210 then :154-180 per request).

Also: requests whose inputs the engine must convert on the caller's stream
(float32, non-contiguous) and that run in the large-request lane give the
same bits through launch / finish as through the synchronous step.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import gp_oracle as O

pytestmark = pytest.mark.gpu

SLICE = list(range(36)) + [880, 1396, 484, 1336]


def _oracle(p):
    """Weight, iteration count, predictive mean and variance of one request:
    Newton in the B-form (equal to the reference's LU form, pinned by
    tests/test_oracle.py) from zeros at tol 1e-5, then the dense-solve predict."""
    kind = O.EQ if p.kind == "EQ" else O.MATERN12
    spec = O.KernelSpec.make(kind, p.variance, p.lengthscale, p.D)
    lik = O.GAUSSIAN if p.lik == "gaussian" else O.PROBIT
    lp = p.likelihood_parameters()
    K = O.gram(spec, p.X)
    w, it, _ = O.laplace_weight(K, p.y, lik, lp, tol=1e-5, form="bform")
    W, _ = O.laplace_precision(K, p.y, lik, lp, w)
    m, v = O.predict(spec, p.X, p.Xtest, w, W)
    return w, it, m, v


@pytest.fixture(scope="module")
def stream_slice():
    from probit_amd import data
    stream = data.serving_stream(seed=4, n=4096)
    reqs = [stream[i] for i in SLICE]
    print(f"c5 slice: N {[r.problem.N for r in reqs]}", flush=True)
    refs = {r.rid: _oracle(r.problem) for r in reqs}
    return reqs, refs


def _serve_async(engine, reqs, torch):
    """Offer every request in order (each refusal ends a round of offers),
    one asynchronous step at a time, until all have completed."""
    dev = [(r.rid, r.problem, torch.as_tensor(r.problem.X, device="cuda"),
            torch.as_tensor(r.problem.y, device="cuda"),
            torch.as_tensor(r.problem.Xtest, device="cuda")) for r in reqs]
    pending, done = list(dev), {}
    while pending or engine.in_flight:
        launched = engine.launch()
        while pending:
            rid, p, X, y, Xt = pending[0]
            if not engine.submit(rid, X, y, p.kind, p.lengthscale, p.variance, p.lik,
                                 p.noise_std, p.cutpoints, Xt, p.period):
                break
            pending.pop(0)
        if launched:
            for c in engine.finish():
                done[c.tag] = c
    return done


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
def test_serving_stream_slice_matches_oracle(stream_slice, precision):
    import torch

    from probit_amd import _lib
    from probit_amd.serving import ServingEngine
    reqs, refs = stream_slice
    eng = ServingEngine(precision=precision, capacity=16)
    print(f"c5 slice {precision}", flush=True)
    done = _serve_async(eng, reqs, torch)
    eng.close()
    assert sorted(done) == sorted(r.rid for r in reqs)
    tol = 1e-9 if precision == "fp64" else 1e-5
    # variance: FP32 / BF16 factors get the FP64 correction of the quadratic
    # form (its error is quadratic in the FP32 solve's)
    vtol = tol
    worst = {"mean": 0.0, "var": 0.0, "extra_iterations": 0}
    for r in reqs:
        c = done[r.rid]
        w_ref, it_ref, m_ref, v_ref = refs[r.rid]
        bad = c.status & ~_lib.INFO_ESCALATED
        assert bad == 0, (r.rid, r.problem.N, c.status)
        if precision == "fp64":
            assert c.iterations == it_ref, (r.rid, c.iterations, it_ref)
        else:
            # an FP32 / BF16 factor with the FP64 residual is an inexact Newton
            # to the same fixed point: the same count or a few more steps (BF16
            # updates contract more slowly until a stall escalates them)
            extra = 4 if precision == "fp32" else 8
            assert it_ref <= c.iterations <= it_ref + extra, (r.rid, c.iterations, it_ref)
        em = np.max(np.abs(c.mean.cpu().numpy() - m_ref)) / np.max(np.abs(m_ref))
        ev = np.max(np.abs(c.var.cpu().numpy() - v_ref)) / np.max(np.abs(v_ref))
        ew = np.max(np.abs(c.weight.cpu().numpy() - w_ref)) / np.max(np.abs(w_ref))
        assert em <= tol and ev <= vtol and ew <= tol, \
            (r.rid, r.problem.N, r.problem.D, r.problem.lik, em, ev, ew)
        worst["mean"] = max(worst["mean"], em)
        worst["var"] = max(worst["var"], ev)
        worst["extra_iterations"] = max(worst["extra_iterations"], c.iterations - it_ref)
    print(f"c5 slice {precision}: worst relative errors {worst}", flush=True)


def test_large_lane_converted_inputs_async_equals_step():
    """Float32 and non-contiguous device inputs are converted on the caller's
    stream; requests with N > 1024 run in the large-request lane, whose steps
    must follow those conversions.  launch / finish gives the bits of the
    synchronous step."""
    import torch

    from probit_amd.serving import ServingEngine
    rng = np.random.default_rng(70)
    probs = []
    for k, N in enumerate((1300, 1900, 1100, 300)):
        D = 3
        X = rng.uniform(-1, 1, (N, D))
        f = np.sin(3 * X.sum(1))
        if k % 2:
            cut = np.array([-np.inf, -0.3, 0.4, np.inf])
            y = np.searchsorted(cut[1:-1], f + 0.4 * rng.normal(size=N)).astype(np.float64)
            probs.append((X, y, "probit", 0.5, cut))
        else:
            probs.append((X, f + 0.2 * rng.normal(size=N), "gaussian", 0.2, None))
    Xt = rng.uniform(-1, 1, (256, 3))

    def run(async_):
        eng = ServingEngine(precision="fp32", capacity=8)
        for tag, (X, y, lik, noise, cut) in enumerate(probs):
            # X as a transposed view of a (D, N) float32 tensor, y in float32:
            # both need a conversion kernel on the caller's stream
            Xd = torch.as_tensor(np.ascontiguousarray(X.T), dtype=torch.float32,
                                 device="cuda").t()
            yd = torch.as_tensor(y, dtype=torch.float32, device="cuda")
            Xtd = torch.as_tensor(Xt.T.copy(), dtype=torch.float32, device="cuda").t()
            assert not Xd.is_contiguous()
            assert eng.submit(tag, Xd, yd, "EQ", 0.7, 1.2, lik, noise, cut, Xtd)
        out = {}
        while eng.in_flight:
            if async_:
                if eng.launch():
                    for c in eng.finish():
                        out[c.tag] = c
            else:
                for c in eng.step():
                    out[c.tag] = c
        eng.close()
        return out

    a, s = run(True), run(False)
    assert sorted(a) == sorted(s) == list(range(len(probs)))
    for t in a:
        for name in ("mean", "var", "weight", "precision"):
            assert torch.equal(getattr(a[t], name), getattr(s[t], name)), (t, name)
        assert a[t].iterations == s[t].iterations and a[t].status == s[t].status == 0


@pytest.mark.parametrize("split_min", ["0", "512"])
def test_size_lanes_async_equals_step(monkeypatch, split_min):
    """Requests in all four size lanes (padded N up to 1024, 1536, 2560 and
    above), which a loaded engine splits the large ones over (here from the
    first member on, PGP_LANE_SPLIT_MIN=0) and a lightly loaded one keeps in
    the second lane (the default, 512 members in flight): launch / finish
    gives the bits of the synchronous step, which has no lanes."""
    import torch

    from probit_amd.serving import ServingEngine
    monkeypatch.setenv("PGP_LANE_SPLIT_MIN", split_min)
    rng = np.random.default_rng(71)
    probs = []
    for k, N in enumerate((300, 1200, 2000, 3000, 800, 1500, 2600, 100)):
        X = rng.uniform(-1, 1, (N, 2))
        f = np.sin(3 * X.sum(1))
        Xt = rng.uniform(-1, 1, (256, 2))
        if k % 2:
            cut = np.array([-np.inf, -0.3, 0.4, np.inf])
            y = np.searchsorted(cut[1:-1], f + 0.4 * rng.normal(size=N)).astype(np.float64)
            probs.append((X, y, "probit", 0.5, cut, Xt))
        else:
            probs.append((X, f + 0.2 * rng.normal(size=N), "gaussian", 0.2, None, Xt))

    def run(async_):
        eng = ServingEngine(precision="fp32", capacity=8)
        for tag, (X, y, lik, noise, cut, Xt) in enumerate(probs):
            assert eng.submit(tag, X, y, "EQ", 0.6, 1.1, lik, noise, cut, Xt)
        out = {}
        while eng.in_flight:
            if async_:
                if eng.launch():
                    for c in eng.finish():
                        out[c.tag] = c
            else:
                for c in eng.step():
                    out[c.tag] = c
        eng.close()
        return out

    a, s = run(True), run(False)
    assert sorted(a) == sorted(s) == list(range(len(probs)))
    for t in a:
        for name in ("mean", "var", "weight", "precision"):
            assert torch.equal(getattr(a[t], name), getattr(s[t], name)), (t, name)
        assert a[t].iterations == s[t].iterations and a[t].status == s[t].status == 0
