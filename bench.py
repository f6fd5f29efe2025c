#!/usr/bin/env python3
"""GP posteriors/sec of the batched mixed-precision serving path on one MI355X.

Workload: the serving request stream (seeded synthetic data): N log-uniform
on [64, 4096], D uniform on {1..8}, X ~ U(-1, 1)^D, Gaussian (noise std 0.2)
or ordinal probit (J uniform on {2..5}), EQ or Matern12 with lengthscale
log-uniform on [0.2, 2], 256 test points per request.  One posterior =
LaplaceGP.approximate_posterior + predict for one request.

The timed region serves --steps x --requests-per-step requests of the stream
as one continuous stream through the continuous-batching engine (closed
loop: every request is offered at the start; requests enter as slots and
memory free up, so the engine stays full until the stream runs out); a step
is --requests-per-step of them (ms_per_step = elapsed / steps).  Closed-loop
latency is measured from a request's admission to its completion.  All
request inputs are resident in HBM before timing starts.  Also measured,
outside the timed region:
  * the plain per-request FP64 path (pgp_laplace_fit + pgp_predict, one
    request at a time) on the same requests: every distinct request of the
    timed set, its time weighted by how often the set holds it (the set
    cycles the stream), and the maximum relative predictive-mean error of the
    batched path against it;
  * per-class launch times (outer update, per-column steps, TRSV, Gram,
    TRSM) over a strided window of the timed requests served again with every
    launch timer on;
  * open-loop latency (the timed requests with Poisson arrivals at 0.7 x the
    measured throughput, from arrival to completion) p50 / p99;
  * the trailing-update kernel's roofline from HIP events over the timed
    region (pgp_timer_*), and the CPU oracle baseline on a stratified sample
    of the timed requests (one per equal-count N-stratum).
--cpu-study REPEATS times only the CPU baseline plan (both oracle variants,
all cores and one core, medians).  Prints ONE JSON line.
"""

from __future__ import annotations

import argparse
import json
import math
import os

# The serving engine runs each size lane on two streams of its own beside the
# caller's; the HIP runtime maps a process's streams onto GPU_MAX_HW_QUEUES
# hardware queues (4 by default), and streams that share a queue serialise.
# The runtime reads it when it starts, so it is raised to 12 (four lanes'
# eight streams, the caller's and torch's own) before torch is imported.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 12:
    os.environ["GPU_MAX_HW_QUEUES"] = "12"
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"bf16": 2500.0, "fp32": 157.3, "fp64": 78.6}
HBM_GBPS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=4)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--requests-per-step", type=int, default=512)
    p.add_argument("--precision", default="fp32", choices=["bf16", "fp32", "fp64"])
    p.add_argument("--capacity", type=int, default=2048,
                   help="request slots of the serving engine (requests in flight)")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="only copy a MIB-sized device buffer 3 times and exit (the known-size "
                        "reference of the HBM traffic counters)")
    p.add_argument("--refine", type=int, default=-1,
                   help="FP64 refinement steps per Newton solve (-1: library default)")
    p.add_argument("--plain-sample", type=int, default=0,
                   help="requests of the plain FP64 leg: 0 = every distinct request of the "
                        "timed set (weighted by how often the set holds it), else a strided "
                        "sample of that many")
    p.add_argument("--profile-window", type=int, default=1024,
                   help="timed requests (strided) served again with every launch-timer class "
                        "on, for the per-class time shares (0: skip)")
    p.add_argument("--open-loop", type=int, default=4096,
                   help="requests of the open-loop (Poisson arrivals) latency run")
    p.add_argument("--rate", type=float, default=0.0,
                   help="open-loop arrival rate (req/s); 0 = 0.7 x measured throughput")
    p.add_argument("--cpu-strata", type=int, default=48,
                   help="CPU baseline: requests timed, one per equal-count N-stratum")
    p.add_argument("--cpu-study", type=int, default=0, metavar="REPEATS",
                   help="only time the CPU baseline plan (both oracle variants, all cores and "
                        "one, median of REPEATS runs) and print it")
    p.add_argument("--seed", type=int, default=4)
    p.add_argument("--stream", type=int, default=4096)
    p.add_argument("--kernel-window", type=int, default=0,
                   help="only serve this many requests of the timed set with synchronous steps "
                        "(the same launches on every run: the command the per-kernel profiles "
                        "and counter runs use) and exit")
    p.add_argument("--factor-only", type=int, default=0, metavar="N",
                   help="only factor one B of an N-point ordinal-probit problem with FP32 and "
                        "BF16 updates (twice each, pgp_laplace_set_state) and print the outer "
                        "update's launch-timer rate")
    p.add_argument("--min-n", type=int, default=0,
                   help="serve only the stream's requests with at least this many points "
                        "(0: the whole stream)")
    return p.parse_args()


def upload(reqs, torch):
    out = []
    for r in reqs:
        p = r.problem
        out.append(dict(rid=r.rid, arrival=r.arrival, p=p,
                        X=torch.as_tensor(p.X, device="cuda"),
                        y=torch.as_tensor(p.y, device="cuda"),
                        Xt=torch.as_tensor(p.Xtest, device="cuda")))
    return out


def serve_window(engine, window, t_arrive=None, chunk=128):
    """Serve a list of requests to completion; returns {rid: (latency_s,
    Completed)}, latency from arrival (t_arrive) or else from submission.
    Asynchronous steps: while the GPU runs a step, the host submits up to
    `chunk` more requests (admitted at the next launch)."""
    import time as _t
    pending = list(window)
    t_sub = {}
    res = {}
    while pending or engine.in_flight:
        launched = engine.launch()
        # requests are offered in order; the first one refused (no slot or no
        # HBM), or a lane's step being done, ends this round of submissions
        n = 0
        while n < len(pending) and n < chunk:
            if n % 8 == 7 and engine.ready():
                break
            r = pending[n]
            if t_arrive is not None and t_arrive[r["rid"]] > _t.perf_counter():
                break
            p = r["p"]
            if not engine.submit(r["rid"], r["X"], r["y"], p.kind, p.lengthscale, p.variance,
                                 p.lik, p.noise_std, p.cutpoints, r["Xt"], p.period):
                break
            t_sub[r["rid"]] = _t.perf_counter()
            n += 1
        pending = pending[n:]
        if not launched:
            continue
        for c in engine.finish():
            t_done = _t.perf_counter()
            start = t_sub[c.tag] if t_arrive is None else t_arrive[c.tag]
            res[c.tag] = (t_done - start, c)
    return res


def kernel_window(engine, window):
    """Serve `window` with synchronous steps, offering requests in order as
    slots free up: the launch sequence depends only on the requests, not on
    timing.  Returns {rid: Completed}."""
    pending, done = list(window), {}
    while pending or engine.in_flight:
        while pending:
            r = pending[0]
            p = r["p"]
            if not engine.submit(r["rid"], r["X"], r["y"], p.kind, p.lengthscale, p.variance,
                                 p.lik, p.noise_std, p.cutpoints, r["Xt"], p.period):
                break
            pending.pop(0)
        for c in engine.step():
            done[c.tag] = c
    return done


def plain_fp64(sample, torch, _lib):
    """One request at a time through the single-problem entry points (FP64).
    Returns {sid: (mean, var, iterations)} and {sid: seconds}."""
    from probit_amd.kernels import EQ, Matern12
    from probit_amd.utilities import likelihood_struct
    lib = _lib.lib()
    stream = _lib.stream_handle()
    out, secs = {}, {}
    torch.cuda.synchronize()
    for r in sample:
        ts = time.perf_counter()
        p = r["p"]
        N, D = p.X.shape
        M = p.Xtest.shape[0]
        base = EQ() if p.kind == "EQ" else Matern12()
        ks = (p.variance * base.stretch(p.lengthscale)).to_struct(D)
        kl = _lib.LIK_GAUSSIAN if p.lik == "gaussian" else _lib.LIK_PROBIT
        ls = likelihood_struct(kl, p.likelihood_parameters())
        ws = _lib.workspace(lib.pgp_workspace_size(N, D, 0), "cuda")
        pws = _lib.workspace(lib.pgp_predict_workspace_size(N, M, 0), "cuda")
        w = torch.empty(N, dtype=torch.float64, device="cuda")
        prec = torch.empty_like(w)
        mean = torch.empty(M, dtype=torch.float64, device="cuda")
        var = torch.empty_like(mean)
        prob = _lib.ProblemT(N, D, 0, r["X"].data_ptr(), r["y"].data_ptr(), w.data_ptr(),
                             prec.data_ptr(), None)
        opts = _lib.FitOptionsT(1e-5, 100, 0, 0, 0)
        res = _lib.FitResultT()
        _lib.check(lib.pgp_laplace_fit(prob, ks, ls, opts, res, ws.data_ptr(), ws.numel(),
                                       stream), "fit")
        _lib.info_bits(lib.pgp_predict(ks, r["Xt"].data_ptr(), M, mean.data_ptr(),
                                       var.data_ptr(), ws.data_ptr(), ws.numel(),
                                       pws.data_ptr(), pws.numel(), stream), "predict")
        torch.cuda.synchronize()
        secs[r["sid"]] = time.perf_counter() - ts
        out[r["sid"]] = (mean, var, res.iterations)
    return out, secs


def _cpu_request_seconds(p, variant):
    """One posterior on the host with the oracle: 'lu' = reference-faithful
    (dense Jacobian + LU Newton per step, dense predict solve), 'bform' = the
    Cholesky B-form Newton."""
    from oracle import gp_oracle as O
    t0 = time.perf_counter()
    kind = O.EQ if p.kind == "EQ" else O.MATERN12
    spec = O.KernelSpec.make(kind, p.variance, p.lengthscale, p.D)
    lik = O.GAUSSIAN if p.lik == "gaussian" else O.PROBIT
    lp = p.likelihood_parameters()
    K = O.gram(spec, p.X)
    w, _, _ = O.laplace_weight(K, p.y, lik, lp, tol=1e-5, form=variant)
    W, _ = O.laplace_precision(K, p.y, lik, lp, w)
    O.predict(spec, p.X, p.Xtest, w, W)
    return time.perf_counter() - t0


def _strata(problems, n):
    """n requests standing for n equal-count strata of the given requests
    ordered by N (the median of each stratum): their mean cost estimates the
    mean cost per request of the whole set, whose cost grows like N^3."""
    order = sorted(range(len(problems)), key=lambda i: problems[i].N)
    n = max(1, min(n, len(order)))
    edges = [round(k * len(order) / n) for k in range(n + 1)]
    return [problems[order[(edges[k] + edges[k + 1] - 1) // 2]] for k in range(n)]


def cpu_baseline(problems, strata, variant="lu", threads=None):
    """The oracle on the host over a stratified sample of the timed requests
    (one request per equal-count N-stratum): posteriors/s = 1 / mean seconds."""
    try:
        from threadpoolctl import threadpool_info, threadpool_limits
    except Exception:  # pragma: no cover
        threadpool_info = threadpool_limits = None
    ctx = threadpool_limits(limits=threads) if (threadpool_limits and threads) else None
    try:
        if ctx is not None:
            ctx.__enter__()
        cores = (max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
                 if threadpool_info else (os.cpu_count() or 1))
        sample = _strata(problems, strata)
        t = [_cpu_request_seconds(p, variant) for p in sample]
    finally:
        if ctx is not None:
            ctx.__exit__(None, None, None)
    return sample, t, int(cores)


def cpu_study(problems, strata, repeats):
    """Both oracle variants, all host cores and one, median of `repeats` runs
    of each stratified sample (the CPU-baseline plan); one JSON line."""
    import platform
    model = platform.processor()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    out = {"host_cpu": model, "nproc": os.cpu_count(), "strata": strata, "repeats": repeats,
           "runs": {}}
    for variant in ("lu", "bform"):
        for threads in (None, 1):
            rates = []
            key = f"{variant}_{'all' if threads is None else 1}_cores"
            for rep in range(repeats):
                sample, t, cores = cpu_baseline(problems, strata, variant, threads)
                rates.append(len(t) / sum(t))
                print(f"cpu_study {key} run {rep}: {rates[-1]:.4f} posteriors/s", file=sys.stderr,
                      flush=True)
            out["runs"][key] = {"cores": cores, "posteriors_per_s_median": float(np.median(rates)),
                                "runs": [round(r, 4) for r in rates]}
    out["sample_mean_N"] = float(np.mean([p.N for p in _strata(problems, strata)]))
    return out


def _mix(problems):
    n = np.array([p.N for p in problems], dtype=np.float64)
    return {"requests": len(problems), "mean_N": round(float(n.mean()), 1),
            "cbrt_mean_N3": round(float(np.mean(n ** 3) ** (1 / 3)), 1)}


TIMER_CLASSES = (("outer_update", "TIMER_SYRK", "mfma"),
                 ("merged_column_steps", "TIMER_SUPER", "mfma"),
                 ("column_steps", "TIMER_PANEL", "mfma"),
                 ("trsv", "TIMER_TRSV", "hbm"), ("gram", "TIMER_GRAM", "hbm"),
                 ("trsm", "TIMER_TRSM", "mfma"), ("variance_refine", "TIMER_REFINE", "mfma"))


def read_timer(lib, _lib, cls):
    import ctypes as C
    n_l, ms, fl, by = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
    _lib.check(lib.pgp_timer_read(cls, C.byref(n_l), C.byref(ms), C.byref(fl), C.byref(by)),
               "pgp_timer_read")
    return n_l.value, ms.value, fl.value, by.value


def factor_only(N, torch, _lib, lib):
    """The outer update's rate in one large factorisation: B = I + W^1/2 K W^1/2
    of an ordinal-probit problem (D = 3, EQ, a small random weight) factored by
    pgp_laplace_set_state with FP32 and with BF16 updates, twice each, the
    outer-update launch timer on.  One JSON line per factorisation."""
    import ctypes as C

    from probit_amd.utilities import likelihood_struct
    D = 3
    rng = np.random.default_rng(3)
    X = rng.normal(size=(N, D))
    cut = np.array([-np.inf, -0.3, 0.4, np.inf])
    y = np.searchsorted(cut[1:-1], np.sin(X.sum(1)) + 0.5 * rng.normal(size=N)).astype(np.float64)
    a = 0.05 * rng.normal(size=N)
    ks = _lib.KernelT()
    ks.kind, ks.dim, ks.variance = _lib.KERNEL_EQ, D, 1.2
    for d in range(D):
        ks.lengthscale[d] = 0.8
    ls = likelihood_struct(_lib.LIK_PROBIT, (0.5, cut))
    Xd, yd, ad = (torch.as_tensor(v, device="cuda") for v in (X, y, a))
    for prec in (_lib.PREC_FP32, _lib.PREC_BF16, _lib.PREC_FP32, _lib.PREC_BF16):
        W = torch.empty(N, dtype=torch.float64, device="cuda")
        prob = _lib.ProblemT(N, D, 0, Xd.data_ptr(), yd.data_ptr(), None, W.data_ptr(), None)
        ws = _lib.workspace(lib.pgp_workspace_size(N, D, prec), "cuda")
        res = _lib.FitResultT()
        torch.cuda.synchronize()
        lib.pgp_timer_reset()
        lib.pgp_timer_enable(1 << _lib.TIMER_SYRK)
        _lib.check(lib.pgp_laplace_set_state(prob, ks, ls, prec, ad.data_ptr(), None, res,
                                             ws.data_ptr(), ws.numel(), _lib.stream_handle()),
                   "pgp_laplace_set_state")
        torch.cuda.synchronize()
        lib.pgp_timer_enable(0)
        n_l, ms, fl, _ = read_timer(lib, _lib, _lib.TIMER_SYRK)
        print(json.dumps({"factor_only": {
            "precision": "fp32" if prec == _lib.PREC_FP32 else "bf16", "N": N,
            "outer_update_launches": n_l, "avg_launch_us": round(1e3 * ms / n_l, 2),
            "algo_flops_per_launch": fl / n_l,
            "tflops": round(fl / (ms * 1e-3) / 1e12, 2), "log_det": res.log_det}}), flush=True)
        del ws


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (got --gpus {args.gpus}); the "
              "serving path shards by routing requests to independent per-GPU engines, "
              "which is not run here", file=sys.stderr)
        sys.exit(2)
    from probit_amd import data
    R, K, W = args.requests_per_step, args.steps, args.warmup
    total = (K + W) * R
    if args.cpu_study:
        stream = data.serving_stream(seed=args.seed, n=min(args.stream, total))
        timed = [stream[i % len(stream)].problem for i in range(W * R, total)]
        out = cpu_study(timed, args.cpu_strata, args.cpu_study)
        out["timed_set"] = _mix(timed)
        print(json.dumps({"cpu_study": out}))
        return
    import torch

    from probit_amd import _lib
    from probit_amd.serving import ServingEngine
    assert torch.cuda.is_available(), "bench.py needs the GPU"
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read_per_copy": 4 * n,
                                               "bytes_written_per_copy": 4 * n, "copies": 3}}))
        return
    lib = _lib.lib()
    if args.factor_only:
        factor_only(args.factor_only, torch, _lib, lib)
        return

    stream = data.serving_stream(seed=args.seed, n=min(args.stream, total))
    if args.min_n > 0:
        stream = [r for r in data.serving_stream(seed=args.seed, n=args.stream)
                  if r.problem.N >= args.min_n]
        if not stream:
            sys.exit(f"bench.py: no request of the stream has N >= {args.min_n}")
    order = [stream[i % len(stream)] for i in range(total)]
    dev_reqs = upload(order, torch)
    # distinct request ids even when the stream is cycled; sid = stream index
    for i, r in enumerate(dev_reqs):
        r["rid"] = i
        r["sid"] = i % len(stream)
    torch.cuda.synchronize()

    eng = ServingEngine(precision=args.precision, capacity=args.capacity,
                        refine=None if args.refine < 0 else args.refine)
    if args.kernel_window:
        # every launch-timer class on: HIP events around the launches (no
        # kernels), so a profiler sees the same launches with or without them
        win = dev_reqs[W * R:W * R + args.kernel_window]
        lib.pgp_timer_reset()
        mask = 0
        for _, attr, _ in TIMER_CLASSES:
            mask |= 1 << getattr(_lib, attr)
        lib.pgp_timer_enable(mask)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = kernel_window(eng, win)
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        lib.pgp_timer_enable(0)
        classes = {}
        for name, attr, _ in TIMER_CLASSES:
            n_l, ms, fl, by = read_timer(lib, _lib, getattr(_lib, attr))
            if n_l:
                classes[name] = {"launches": n_l, "avg_launch_us": round(1e3 * ms / n_l, 3),
                                 "algo_flops_per_launch": fl / n_l,
                                 "algo_bytes_per_launch": by / n_l}
        print(json.dumps({"kernel_window": {"requests": len(done), "seconds": round(secs, 4),
                                            "precision": args.precision,
                                            "timed_set": _mix([r["p"] for r in win]),
                                            "timer_classes": classes}}))
        return
    # warmup windows
    for s in range(W):
        serve_window(eng, dev_reqs[s * R:(s + 1) * R])
    torch.cuda.synchronize()
    lib.pgp_timer_reset()
    lib.pgp_timer_enable((1 << _lib.TIMER_SUPER) | (1 << _lib.TIMER_SYRK))
    timed = dev_reqs[W * R:(W + K) * R]
    results = {}
    hs0 = eng.host_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    results.update(serve_window(eng, timed))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    hs1 = eng.host_stats()
    lib.pgp_timer_enable(0)
    timed_classes = {c: read_timer(lib, _lib, c) for c in (_lib.TIMER_SUPER, _lib.TIMER_SYRK)}
    elapsed = t1 - t0
    value = K * R / elapsed
    lat_closed = np.array([v[0] for v in results.values()])
    iters = np.array([v[1].iterations for v in results.values()])

    # per-class launch times over a window of the timed requests (every timer
    # class on: the events add host work, so this window is not the timed one)
    shares = None
    if args.profile_window > 0:
        stride = max(1, len(timed) // args.profile_window)
        win = timed[::stride][:args.profile_window]
        for r in win:
            r["rid"] += 10 ** 7
        lib.pgp_timer_reset()
        mask = 0
        for _, attr, _ in TIMER_CLASSES:
            mask |= 1 << getattr(_lib, attr)
        torch.cuda.synchronize()
        lib.pgp_timer_enable(mask)
        tp0 = time.perf_counter()
        serve_window(eng, win)
        torch.cuda.synchronize()
        tp1 = time.perf_counter()
        lib.pgp_timer_enable(0)
        for r in win:
            r["rid"] -= 10 ** 7
        shares = {"requests": len(win), "wall_ms": round(1e3 * (tp1 - tp0), 3), "classes": {}}
        for name, attr, bound in TIMER_CLASSES:
            n, t_ms, f_, b_ = read_timer(lib, _lib, getattr(_lib, attr))
            shares["classes"][name] = {
                "launches": n, "total_ms": round(t_ms, 3),
                "avg_launch_us": round(1e3 * t_ms / max(1, n), 3),
                "share_of_wall": round(t_ms / max(1e-9, 1e3 * (tp1 - tp0)), 4),
                "tflops": round(f_ / max(1e-12, t_ms * 1e-3) / 1e12, 3) if f_ else None,
                "gbps": round(b_ / max(1e-12, t_ms * 1e-3) / 1e9, 3) if b_ else None,
            }

    # plain FP64 per-request path on the same requests: every distinct
    # request of the timed set, its time weighted by how often the set holds it
    sids = [r["sid"] for r in timed]
    mult = {}
    for sid in sids:
        mult[sid] = mult.get(sid, 0) + 1
    first = {}
    for r in timed:
        first.setdefault(r["sid"], r)
    distinct = [first[k] for k in sorted(first)]
    if args.plain_sample > 0:
        step = max(1, len(distinct) // args.plain_sample)
        distinct = distinct[::step][:args.plain_sample]
    plain, secs = plain_fp64(distinct, torch, _lib)
    w_tot = sum(mult[r["sid"]] for r in distinct)
    plain_rate = w_tot / sum(mult[r["sid"]] * secs[r["sid"]] for r in distinct)
    plain_lat = np.repeat([secs[r["sid"]] for r in distinct], [mult[r["sid"]] for r in distinct])
    errs = []
    by_sid = {}
    for rid, (_, c) in results.items():
        by_sid.setdefault(dev_reqs[rid]["sid"], c)
    for r in distinct:
        c = by_sid[r["sid"]]
        m_ref = plain[r["sid"]][0].cpu().numpy()
        m = c.mean.cpu().numpy()
        errs.append(float(np.max(np.abs(m - m_ref)) / max(1e-300, np.max(np.abs(m_ref)))))
    errs = np.array(errs)

    # open-loop latency at a fixed arrival rate: the stream's requests with
    # Poisson arrivals, latency from arrival to completion
    rate = args.rate if args.rate > 0 else 0.7 * value
    ol = timed[:min(args.open_loop, len(timed))]
    for r in ol:
        r["rid"] += 2 * 10 ** 7
    t_base = time.perf_counter() + 0.01
    gaps = np.random.default_rng(args.seed + 1).exponential(1.0 / rate, len(ol))
    arrive = {r["rid"]: t_base + float(t) for r, t in zip(ol, np.cumsum(gaps))}
    open_res = serve_window(eng, ol, t_arrive=arrive)
    for r in ol:
        r["rid"] -= 2 * 10 ** 7
    lat_open = np.array([v[0] for v in open_res.values()])

    timed_problems = [r["p"] for r in timed]
    sample, t_cpu, cores = cpu_baseline(timed_problems, args.cpu_strata)
    cpu = dict(value=len(t_cpu) / sum(t_cpu), unit="posteriors/s", cores=cores, kind="port",
               sample=f"{len(sample)} requests, one per equal-count N-stratum of the timed set "
                      f"(N {min(p.N for p in sample)}..{max(p.N for p in sample)}); "
                      f"reference-faithful LU Newton + dense predict solve (the oracle), NumPy "
                      f"FP64, {sum(t_cpu):.1f} s; posteriors/s = requests / total seconds")

    # Rooflines over the timed region from the library's HIP-event timers:
    # the dominant kernel (the merged column kernel of the Cholesky,
    # k_superstep) and the outer trailing update.  achieved = algorithmic
    # flops per launch / average launch duration; traffic = this run's
    # algorithmic bytes per launch x the HBM bytes per algorithmic byte that
    # the committed FETCH_SIZE / WRITE_SIZE runs (copy-calibrated) measured
    # on the same kernel (profiles/roofline_traffic.json; the launch mix
    # varies with the serving dynamics), null if not measured.
    peak = PEAK_TFLOPS[args.precision]
    tpath = os.path.join(ROOT, "profiles", "roofline_traffic.json")
    ratios = {}
    if os.path.exists(tpath):
        with open(tpath) as f:
            ratios = json.load(f).get(args.precision, {})

    def roofline(cls, kernel, label):
        n_l, ms, fl, by = timed_classes[cls]
        achieved = fl / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
        algo_bytes = by / max(1, n_l)
        rec = ratios.get(kernel) or {}
        traffic = (round(rec["hbm_bytes_per_algo_byte"] * algo_bytes, -3)
                   if rec.get("hbm_bytes_per_algo_byte") else None)
        return {"kernel": f"{kernel} ({label})", "bound": "mfma", "achieved": round(achieved, 3),
                "peak": peak, "unit": "TFLOP/s", "frac": round(achieved / peak, 5),
                "traffic": traffic, "launches": int(n_l),
                "avg_launch_us": round(1e3 * ms / max(1, n_l), 3),
                "algo_flops_per_launch": fl / max(1, n_l), "algo_bytes_per_launch": algo_bytes}

    line = {
        "metric": "GP posteriors/sec (batched mixed-precision serving, p99 latency reported)",
        "value": round(value, 2),
        "unit": "posteriors/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 * elapsed / K, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16": "bf16-mfma+fp32-acc+fp64", "fp32": "fp32-mfma+fp64",
                  "fp64": "fp64"}[args.precision],
        "data": "synthetic",
        "config": {
            "workload": "serving stream: N logU[64,4096], D U{1..8}, Gaussian / probit J U{2..5}, "
                        "EQ / Matern12, 256 test points; approximate_posterior + predict",
            "requests_per_step": R,
            "capacity": args.capacity,
            "lane_np": os.environ.get("PGP_LANE_NP", "1024,1536,2560"),
            "lane_split_min": int(os.environ.get("PGP_LANE_SPLIT_MIN", "512")),
            "hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"]),
            "min_n": args.min_n,
            "precision": args.precision,
            "tolerance": 1e-5,
            "timed_set": _mix(timed_problems),
        },
        "p50_latency_ms": round(1e3 * float(np.percentile(lat_open, 50)), 3),
        "p99_latency_ms": round(1e3 * float(np.percentile(lat_open, 99)), 3),
        "open_loop": {
            "rate_req_per_s": round(rate, 1),
            "requests": len(lat_open),
            "p50_latency_ms": round(1e3 * float(np.percentile(lat_open, 50)), 3),
            "p99_latency_ms": round(1e3 * float(np.percentile(lat_open, 99)), 3),
            "latency_from": "arrival",
        },
        "hbm_in_flight": {"peak_gb": round(eng.peak_used / 1e9, 2),
                          "budget_gb": round(eng.budget / 1e9, 2)},
        "closed_loop": {
            "p50_latency_ms": round(1e3 * float(np.percentile(lat_closed, 50)), 3),
            "p99_latency_ms": round(1e3 * float(np.percentile(lat_closed, 99)), 3),
            "latency_from": "admission",
        },
        "plain_fp64": {
            "value": round(plain_rate, 2),
            "unit": "posteriors/s",
            "requests": "every distinct request of the timed set, weighted by its multiplicity"
                        if args.plain_sample <= 0 else f"strided sample of {len(distinct)}",
            "distinct_requests": len(distinct),
            "p50_latency_ms": round(1e3 * float(np.percentile(plain_lat, 50)), 3),
            "p99_latency_ms": round(1e3 * float(np.percentile(plain_lat, 99)), 3),
        },
        "speedup_vs_plain_fp64": round(value / plain_rate, 2),
        "max_rel_mean_err_vs_fp64": float(np.max(errs)),
        "mean_newton_iterations": round(float(np.mean(iters)), 2),
        "roofline": roofline(_lib.TIMER_SUPER, "k_superstep",
                             f"{args.precision} merged column-form super-steps of the "
                             "Cholesky: left-looking tile updates, FP64 diagonal factors, panels"),
        "roofline_outer_update": roofline(_lib.TIMER_SYRK, "k_syrk_outer",
                                          f"{args.precision} rank-64G trailing update of the "
                                          "Cholesky, G from the schedule table"),
        "kernel_classes": shares,
        "host_loop": {  # the serving loop's host thread over the timed region
            "lane_steps": hs1["lane_steps"] - hs0["lane_steps"],
            "enqueue_share": round((hs1["enqueue_s"] - hs0["enqueue_s"]) / elapsed, 4),
            "wait_for_gpu_share": round((hs1["wait_s"] - hs0["wait_s"]) / elapsed, 4),
            "report_share": round((hs1["report_s"] - hs0["report_s"]) / elapsed, 4),
            "enqueue_us_per_lane_step": round(1e6 * (hs1["enqueue_s"] - hs0["enqueue_s"]) /
                                              max(1, hs1["lane_steps"] - hs0["lane_steps"]), 1),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
