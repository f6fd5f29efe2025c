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
from .kernels import EQ, Matern12
from .utilities import likelihood_struct


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

    # -- sizing ---------------------------------------------------------------
    def request_bytes(self, N: int, D: int, M: int) -> int:
        b = self.lib.pgp_workspace_size(N, D, self.prec)
        if M > 0:
            b += self.lib.pgp_predict_workspace_size(N, M, self.prec)
        return int(b + 8 * (N * (D + 3) + M * (D + 2)))

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
        Device tensors are used in place; host arrays are copied to the GPU."""
        N, D = tuple(X.shape)
        M = 0 if Xtest is None else int(Xtest.shape[0])
        if not self.can_admit(N, D, M):
            return False
        dev = self.device
        t = {}
        t["X"] = self._dev(X)
        t["y"] = self._dev(y)
        t["w"] = torch.empty(N, dtype=torch.float64, device=dev)
        t["p"] = torch.empty(N, dtype=torch.float64, device=dev)
        nb = self.lib.pgp_workspace_size(N, D, self.prec)
        t["ws"] = _lib.workspace(nb, dev)
        item = _lib.BatchItemT()
        item.problem.N = N
        item.problem.D = D
        item.problem.X = t["X"].data_ptr()
        item.problem.y = t["y"].data_ptr()
        item.problem.weight = t["w"].data_ptr()
        item.problem.precision = t["p"].data_ptr()
        base = EQ() if kind in ("EQ", 0) else Matern12()
        k = variance * base.stretch(lengthscale)
        if period is not None:
            k = k.periodic(period)
        item.kernel = k.to_struct(D)
        kind_l = _lib.LIK_GAUSSIAN if lik == "gaussian" else _lib.LIK_PROBIT
        lp = (noise_std,) if lik == "gaussian" else (noise_std, cutpoints)
        item.lik = likelihood_struct(kind_l, lp)
        item.workspace = t["ws"].data_ptr()
        item.workspace_bytes = nb
        if M > 0:
            t["Xt"] = self._dev(Xtest)
            t["mean"] = torch.empty(M, dtype=torch.float64, device=dev)
            t["var"] = torch.empty(M, dtype=torch.float64, device=dev)
            pb = self.lib.pgp_predict_workspace_size(N, M, self.prec)
            t["pws"] = _lib.workspace(pb, dev)
            item.Xtest = t["Xt"].data_ptr()
            item.M = M
            item.mean = t["mean"].data_ptr()
            item.var = t["var"].data_ptr()
            item.predict_workspace = t["pws"].data_ptr()
            item.predict_workspace_bytes = pb
        rc = self.lib.pgp_batcher_submit(self._h, item, int(tag))
        _lib.check(rc, "pgp_batcher_submit")
        s = _Slot()
        s.item, s.tensors = item, t
        s.nbytes = self.request_bytes(N, D, M)
        s.t_submit = time.perf_counter()
        self._slots[int(tag)] = s
        self.used += s.nbytes
        return True

    def step(self) -> list[Completed]:
        """One Newton iteration for everything in flight; returns finished requests."""
        n = C.c_int32(0)
        rc = self.lib.pgp_batcher_step(self._h, self._done_buf, self.capacity, C.byref(n))
        _lib.check(rc, "pgp_batcher_step")
        out = []
        for i in range(n.value):
            tag = int(self._done_buf[i])
            s = self._slots.pop(tag)
            r = s.item.result
            t = s.tensors
            out.append(Completed(tag, t.get("mean"), t.get("var"), t["w"], t["p"],
                                 r.iterations, r.status, r.objective))
            self.used -= s.nbytes
        return out

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
