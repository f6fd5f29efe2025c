"""Continuous-batching request loop for many independent GP fits.

Each request is one posterior: approximate_posterior + predict at its own
test points (probit/approximators.py:204-210 then :154-180).  Requests are
admitted while the device-memory budget (their workspaces) and the slot
capacity allow; every ``step()`` runs one Newton iteration for all requests
in flight inside shared launches (libprobit_amd's batcher), finalises and
predicts the ones that converged and releases their memory, so requests that
arrive while others are on the GPU join at the next step.
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib


@dataclass
class Completed:
    tag: int
    mean: torch.Tensor
    var: torch.Tensor
    weight: torch.Tensor
    precision: torch.Tensor
    iterations: int
    status: int
    objective: float


class _Slot:
    __slots__ = ("item", "tensors", "nbytes", "t_submit")


class ServingEngine:
    """One process, one GPU: a batcher plus an HBM-budget admitter."""

    def __init__(self, precision="bf16", capacity=256, hbm_budget_bytes=None, tolerance=1e-5,
                 maxiter=100, refine=None, device="cuda"):
        self.lib = _lib.lib()
        self.device = torch.device(device)
        self.prec = _lib.PRECISIONS[precision]
        self.capacity = int(capacity)
        if hbm_budget_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget_bytes = int(free * 0.8)
        self.budget = int(hbm_budget_bytes)
        self.used = 0
        self.peak_used = 0  # the most bytes of admitted requests held at once
        opts = _lib.FitOptionsT()
        opts.tolerance = tolerance
        opts.maxiter = maxiter
        opts.precision = self.prec
        opts.refine = -1 if refine is None else int(refine)
        self._opts = opts
        tb = self.lib.pgp_batcher_table_size(self.capacity)
        self._table = _lib.workspace(tb, self.device)
        h = C.c_void_p()
        rc = self.lib.pgp_batcher_create(self.capacity, opts, self._table.data_ptr(), tb,
                                         _lib.stream_handle(self.device), C.byref(h))
        _lib.check(rc, "pgp_batcher_create")
        self._h = h
        self._slots: dict[int, _Slot] = {}
        self._done_buf = (C.c_int64 * self.capacity)()
        self._launched = False

    # -- sizing ---------------------------------------------------------------
    def _sizes(self, N: int, D: int, M: int):
        """(workspace bytes, predict workspace bytes, request bytes)."""
        nb = int(self.lib.pgp_workspace_size(N, D, self.prec))
        pb = int(self.lib.pgp_predict_workspace_size(N, M, self.prec)) if M > 0 else 0
        return nb, pb, nb + pb + 8 * (N * (D + 3) + M * (D + 2))

    def request_bytes(self, N: int, D: int, M: int) -> int:
        return self._sizes(N, D, M)[2]

    def can_admit(self, N: int, D: int, M: int) -> bool:
        return (len(self._slots) < self.capacity
                and self.used + self.request_bytes(N, D, M) <= self.budget)

    # -- requests ---------------------------------------------------------------
    def _dev(self, a) -> torch.Tensor:
        if isinstance(a, torch.Tensor) and a.device == self.device and a.dtype == torch.float64:
            return a.contiguous()
        if isinstance(a, torch.Tensor):
            return a.to(device=self.device, dtype=torch.float64).contiguous()
        return torch.as_tensor(np.asarray(a, dtype=np.float64), device=self.device)

    def submit(self, tag: int, X, y, kind: str, lengthscale, variance, lik: str, noise_std,
               cutpoints=None, Xtest=None, period=None) -> bool:
        """Queue one posterior; returns False (nothing queued) if it does not fit now.
        Device tensors are used in place; host arrays are copied to the GPU.
        Per request, one device block holds its workspaces (released when the
        request completes, so the allocator's cached blocks serve the next
        requests) and a small one its outputs (what the completion's tensors
        view)."""
        if len(self._slots) >= self.capacity:  # no slot: nothing else to check
            return False
        N, D = int(X.shape[0]), int(X.shape[1])
        M = 0 if Xtest is None else int(Xtest.shape[0])
        nb, pb, nbytes = self._sizes(N, D, M)
        if self.used + nbytes > self.budget:
            return False
        if D > _lib.MAX_DIM:
            raise ValueError(f"input dimension must be in [1, {_lib.MAX_DIM}] (got {D})")
        vec = (8 * N + 255) // 256 * 256
        mv = (8 * M + 255) // 256 * 256
        # workspaces [workspace | predict workspace]; outputs [weight |
        # precision | mean | var]
        o_pw = (nb + 255) // 256 * 256
        ws = torch.empty(o_pw + (pb + 255) // 256 * 256, dtype=torch.uint8, device=self.device)
        o_w = 0
        o_p = o_w + vec
        o_m = o_p + vec
        o_v = o_m + mv
        outb = torch.empty(o_v + mv, dtype=torch.uint8, device=self.device)
        wbase, base = ws.data_ptr(), outb.data_ptr()
        Xd, yd = self._dev(X), self._dev(y)
        item = _lib.BatchItemT()
        pr = item.problem
        pr.N, pr.D = N, D
        pr.X, pr.y = Xd.data_ptr(), yd.data_ptr()
        pr.weight, pr.precision = base + o_w, base + o_p
        k = item.kernel
        k.kind = _lib.KERNEL_EQ if kind in ("EQ", 0) else _lib.KERNEL_MATERN12
        k.dim = D
        k.variance = float(variance)
        ls = np.broadcast_to(np.asarray(lengthscale, dtype=np.float64), (D,))
        k.lengthscale[:D] = ls.tolist()
        if period is not None:
            k.period[:D] = np.broadcast_to(np.asarray(period, dtype=np.float64), (D,)).tolist()
        lk = item.lik
        lk.noise_std = float(noise_std)
        lk.single_precision_bounds = 1
        if lik == "gaussian":
            lk.kind = _lib.LIK_GAUSSIAN
        else:
            cut = np.asarray(cutpoints, dtype=np.float64).reshape(-1)
            if cut.shape[0] < 3 or cut.shape[0] > _lib.MAX_CUT:
                raise ValueError(f"need 3..{_lib.MAX_CUT} cutpoints incl. +-inf")
            lk.kind = _lib.LIK_PROBIT
            lk.J = cut.shape[0] - 1
            lk.cutpoints[:cut.shape[0]] = cut.tolist()
        item.workspace = wbase
        item.workspace_bytes = nb
        keep = [outb, ws, Xd, yd]
        if M > 0:
            Xt = self._dev(Xtest)
            keep.append(Xt)
            item.Xtest = Xt.data_ptr()
            item.M = M
            item.mean, item.var = base + o_m, base + o_v
            item.predict_workspace = wbase + o_pw
            item.predict_workspace_bytes = pb
        rc = self.lib.pgp_batcher_submit(self._h, item, int(tag))
        _lib.check(rc, "pgp_batcher_submit")
        s = _Slot()
        s.item, s.tensors = item, (keep, N, M, o_w, o_p, o_m, o_v)
        s.nbytes = nbytes
        s.t_submit = time.perf_counter()
        self._slots[int(tag)] = s
        self.used += s.nbytes
        self.peak_used = max(self.peak_used, self.used)
        return True

    @staticmethod
    def _outputs(t):
        keep, N, M, o_w, o_p, o_m, o_v = t
        block = keep[0]

        def f64(off, n):
            return block[off:off + 8 * n].view(torch.float64)

        mean = f64(o_m, M) if M > 0 else None
        var = f64(o_v, M) if M > 0 else None
        return f64(o_w, N), f64(o_p, N), mean, var

    def step(self) -> list[Completed]:
        """One Newton iteration for everything in flight; returns finished requests."""
        if self._launched:
            raise RuntimeError("step() while an asynchronous launch is pending: call finish()")
        n = C.c_int32(0)
        rc = self.lib.pgp_batcher_step(self._h, self._done_buf, self.capacity, C.byref(n))
        _lib.check(rc, "pgp_batcher_step")
        return self._collect(n.value)

    def _collect(self, n):
        out = []
        for i in range(n):
            tag = int(self._done_buf[i])
            s = self._slots.pop(tag)
            r = s.item.result
            w, p, mean, var = self._outputs(s.tensors)
            out.append(Completed(tag, mean, var, w, p, r.iterations, r.status, r.objective))
            self.used -= s.nbytes
        return out

    def launch(self) -> bool:
        """Enqueue one step on every lane (small / large requests) that has
        none in flight and return at once (requests submitted meanwhile join
        at the next launch).  Returns True if any step is in flight."""
        if self._slots:
            _lib.check(self.lib.pgp_batcher_launch(self._h), "pgp_batcher_launch")
        self._launched = self.lib.pgp_batcher_launched(self._h) > 0
        return self._launched

    def finish(self) -> list[Completed]:
        """Wait until a lane's step is done, start the next step of every idle
        lane at once (before the completions are collected on the host, so
        the lane does not wait for that), and return the requests it
        completed."""
        if not self._launched:
            return []
        n = C.c_int32(0)
        rc = self.lib.pgp_batcher_finish(self._h, self._done_buf, self.capacity, C.byref(n))
        _lib.check(rc, "pgp_batcher_finish")
        if len(self._slots) > n.value:  # requests still in flight or waiting
            _lib.check(self.lib.pgp_batcher_launch(self._h), "pgp_batcher_launch")
        self._launched = self.lib.pgp_batcher_launched(self._h) > 0
        return self._collect(n.value)

    def ready(self) -> bool:
        """True if a lane's step is done (finish() would return at once)."""
        return self._launched and self.lib.pgp_batcher_ready(self._h) > 0

    def host_stats(self) -> dict:
        """Host-side seconds since creation: enqueueing lane steps (launch),
        waiting for the GPU in finish(), reporting completed steps, and the
        lane steps launched."""
        e, w, r = C.c_double(), C.c_double(), C.c_double()
        n = C.c_int64()
        _lib.check(self.lib.pgp_batcher_stats(self._h, C.byref(e), C.byref(w), C.byref(r),
                                              C.byref(n)), "pgp_batcher_stats")
        return {"enqueue_s": e.value, "wait_s": w.value, "report_s": r.value,
                "lane_steps": n.value}

    @property
    def in_flight(self) -> int:
        return len(self._slots)

    def close(self):
        if getattr(self, "_h", None) is not None:
            self.lib.pgp_batcher_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve_async(engine: ServingEngine, requests, chunk=128):
    """Serve `requests` -- (tag, X, y, kind, lengthscale, variance, lik,
    noise_std, cutpoints, Xtest, period) tuples with device tensors -- to
    completion with asynchronous steps: while the GPU runs a step the host
    offers the next requests in order (a refusal, or a lane's step being
    done, ends a round of offers).  Returns {tag: Completed}."""
    pending = list(requests)
    done = {}
    while pending or engine.in_flight:
        launched = engine.launch()
        n = 0
        while n < len(pending) and n < chunk:
            if n % 8 == 7 and engine.ready():
                break
            if not engine.submit(*pending[n]):
                break
            n += 1
        pending = pending[n:]
        if launched:
            for c in engine.finish():
                done[c.tag] = c
    return done


def serve_stream(engine: ServingEngine, requests, realtime=True, on_done=None):
    """Drive a request stream with arrival times (seconds from start).

    Returns {rid: (latency_s, Completed)}.  With ``realtime`` a request
    becomes eligible at its arrival time on the host clock; latency runs from
    arrival (or submission, if later admission was blocked) to completion.
    """
    reqs = sorted(requests, key=lambda r: r.arrival)
    t0 = time.perf_counter()
    nxt = 0
    waiting = []
    results = {}
    arrival_abs = {}
    while nxt < len(reqs) or waiting or engine.in_flight:
        now = time.perf_counter() - t0
        while nxt < len(reqs) and (not realtime or reqs[nxt].arrival <= now):
            waiting.append(reqs[nxt])
            arrival_abs[reqs[nxt].rid] = t0 + (reqs[nxt].arrival if realtime else 0.0)
            nxt += 1
        still = []
        for r in waiting:
            p = r.problem
            ok = engine.submit(r.rid, p.X, p.y, p.kind, p.lengthscale, p.variance, p.lik,
                               p.noise_std, p.cutpoints, p.Xtest, p.period)
            if not ok:
                still.append(r)
        waiting = still
        if engine.in_flight == 0:
            if nxt < len(reqs) and realtime:
                time.sleep(max(0.0, min(1e-3, reqs[nxt].arrival - (time.perf_counter() - t0))))
            continue
        for c in engine.step():
            t_done = time.perf_counter()
            results[c.tag] = (t_done - arrival_abs[c.tag], c)
            if on_done is not None:
                on_done(c)
    return results
