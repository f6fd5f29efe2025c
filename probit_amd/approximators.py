"""LaplaceGP / VBGP with the reference's API (probit/approximators.py), on HIP.

Signatures follow the reference:
  * ``__init__(data, prior, log_likelihood, grad_log_likelihood=None,
    hessian_log_likelihood=None, tolerance=1e-5)`` (probit/approximators.py:55-63)
  * ``approximate_posterior(parameters) -> (weight, precision)`` (:204-210)
  * ``predict(X_test, parameters, weight, precision) -> (mean, variance)`` (:154-180)
  * ``predict_covariance(...)`` (:182-197), ``objective()`` (:248-263, :316-330),
    ``value_and_grad()`` (:132-134), ``weight``, ``precision``, ``construct``.

Extra keyword-only arguments choose the factor precision ("fp64", "fp32",
"bf16"), the Newton iteration cap and the number of FP64 refinement steps.
All compute runs in libprobit_amd (HIP, gfx950); torch supplies device
memory, streams and the autograd graph around the evidence.  Numerical
failures surface as NaN (as JAX would produce) plus ``last_result.status``;
with ``strict=True`` they raise ``FloatingPointError`` instead.
"""

from __future__ import annotations

import ctypes as C
import functools
import warnings
from abc import ABC, abstractmethod

import numpy as np
import torch

from . import _lib
from .kernels import Kernel
from .utilities import (InvalidKernel, grad_log_probit_likelihood,
                        hessian_log_probit_likelihood, likelihood_struct,
                        log_gaussian_likelihood, log_probit_likelihood)

__all__ = ["Approximator", "LaplaceGP", "VBGP"]


def _closed_form(fn, base):
    """(True, single_precision) if `fn` is `base` or functools.partial(base,
    single_precision=...), else (False, None)."""
    if fn is base:
        return True, True
    if (isinstance(fn, functools.partial) and fn.func is base and not fn.args
            and set(fn.keywords) <= {"single_precision"}):
        return True, bool(fn.keywords.get("single_precision", True))
    return False, None


def _resolve_likelihood(log_likelihood, grad_log_likelihood, hessian_log_likelihood):
    """Map the reference's likelihood callables to a fused-kernel ID and the
    BOUNDS set of the closed-form derivatives (True: BOUNDS["single"], the
    callables' default; functools.partial(..., single_precision=False) picks
    BOUNDS["double"], as it would in the reference)."""
    if log_likelihood is log_gaussian_likelihood:
        if grad_log_likelihood is not None or hessian_log_likelihood is not None:
            raise NotImplementedError("custom derivatives of the Gaussian likelihood")
        return _lib.LIK_GAUSSIAN, True
    if log_likelihood is log_probit_likelihood:
        if grad_log_likelihood is None and hessian_log_likelihood is None:
            return _lib.LIK_PROBIT, True
        g_ok, g_sp = _closed_form(grad_log_likelihood, grad_log_probit_likelihood)
        h_ok, h_sp = _closed_form(hessian_log_likelihood, hessian_log_probit_likelihood)
        if g_ok and h_ok:
            if g_sp != h_sp:
                raise NotImplementedError("the gradient and Hessian use different BOUNDS sets")
            return _lib.LIK_PROBIT_CLOSED, g_sp
    raise NotImplementedError(
        "no fused HIP kernel for this likelihood: use probit_amd.utilities."
        "log_gaussian_likelihood or log_probit_likelihood (optionally with "
        "grad_/hessian_log_probit_likelihood)")


def _flatten(tree):
    if isinstance(tree, (tuple, list)):
        leaves, specs = [], []
        for t in tree:
            lv, sp = _flatten(t)
            leaves += lv
            specs.append(sp)
        return leaves, (type(tree), specs)
    return [tree], None


def _unflatten(spec, leaves):
    it = iter(leaves)

    def rec(sp):
        if sp is None:
            return next(it)
        typ, subs = sp
        return typ(rec(s) for s in subs)

    return rec(spec)


# status bits whose results are not valid numbers (INFO_ESCALATED only
# reports that BF16 updates were replaced by FP32 ones)
_FAILURES = (("non-positive pivot", _lib.INFO_NOT_PD),
             ("negative precision W", _lib.INFO_NEGATIVE_W),
             ("fixed point not reached within maxiter", _lib.INFO_MAXITER),
             ("a solve's bounded wait ran out", _lib.INFO_STALLED))


def numerical_failure(status: int, objective: float = 0.0) -> str | None:
    """What makes a fit result invalid, or None.  The reference returns NaN
    (JAX) and raises nothing; ``strict=True`` raises with this text."""
    names = [name for name, bit in _FAILURES if status & bit]
    if not np.isfinite(objective):
        names.append("non-finite objective")
    return ", ".join(names) or None


def _refuse_closed_form_gradient(lik_kind):
    """The evidence gradient is not provided for the closed-form (BOUNDS)
    probit derivatives: JAX would differentiate through their piecewise tail
    series (probit/utilities.py:88-192), including the branch selection, and
    no adjoint of those pieces is built here.  The value, fit and predict are."""
    if lik_kind == _lib.LIK_PROBIT_CLOSED:
        raise NotImplementedError(
            "the evidence gradient is not provided with grad_/hessian_log_probit_likelihood "
            "(the closed-form BOUNDS derivatives); use log_probit_likelihood alone, whose "
            "derivatives are exact")


class Approximator(ABC):
    """Base class: data upload, parameter packing, workspaces, predict."""

    @abstractmethod
    def __repr__(self):
        """Name of the approximator."""

    def __init__(self, data, prior, log_likelihood, grad_log_likelihood=None,
                 hessian_log_likelihood=None, tolerance=1e-5, *, precision_mode="fp64",
                 maxiter=100, refine=None, device=None, strict=False):
        self.tolerance = float(tolerance)
        self.strict = bool(strict)
        self.prior = prior
        self.lik_kind, self._single_bounds = _resolve_likelihood(
            log_likelihood, grad_log_likelihood, hessian_log_likelihood)
        self.log_likelihood = log_likelihood
        self.grad_log_likelihood = grad_log_likelihood
        self.hessian_log_likelihood = hessian_log_likelihood
        if precision_mode not in _lib.PRECISIONS:
            raise ValueError(f"precision_mode must be one of {list(_lib.PRECISIONS)}")
        self.precision_mode = precision_mode
        self._prec = _lib.PRECISIONS[precision_mode]
        self.maxiter = int(maxiter)
        self.refine = -1 if refine is None else int(refine)
        X_train, y_train = data
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        X = torch.as_tensor(np.asarray(X_train) if not isinstance(X_train, torch.Tensor)
                            else X_train).detach().to(torch.float64)
        if X.dim() == 1:
            X = X[:, None]
        self.N, self.D = X.shape
        y = torch.as_tensor(np.asarray(y_train) if not isinstance(y_train, torch.Tensor)
                            else y_train).detach().reshape(-1).to(torch.float64)
        if y.shape[0] != self.N:
            raise ValueError("X and y disagree on N")
        self._X = X.to(self.device).contiguous()
        self._y = y.to(self.device).contiguous()
        self.data = data
        self._ws = None
        self._pws = None
        self._gws = None
        self._state = None
        self._generation = 0
        self.last_result = None
        self.last_status = 0

    # ---- packing ----------------------------------------------------------
    def _kernel(self, prior_parameters) -> Kernel:
        k = self.prior(prior_parameters)
        if not isinstance(k, Kernel):
            raise InvalidKernel(k)
        return k

    def _structs(self, parameters):
        kern = self._kernel(parameters[0])
        ks = kern.to_struct(self.D)
        ls = likelihood_struct(self.lik_kind, parameters[1], self._single_bounds)
        return ks, ls

    @staticmethod
    def _key(ks, ls):
        return (ks.kind, ks.variance, tuple(ks.lengthscale), tuple(ks.period), ls.kind,
                ls.noise_std, ls.J, tuple(ls.cutpoints))

    def _problem(self, weight=None, precision=None, latent=None) -> _lib.ProblemT:
        p = _lib.ProblemT()
        p.N = self.N
        p.D = self.D
        p.X = self._X.data_ptr()
        p.y = self._y.data_ptr()
        p.weight = _lib.ptr(weight)
        p.precision = _lib.ptr(precision)
        p.latent = _lib.ptr(latent)
        return p

    def _options(self) -> _lib.FitOptionsT:
        o = _lib.FitOptionsT()
        o.tolerance = self.tolerance
        o.maxiter = self.maxiter
        o.precision = self._prec
        o.refine = self.refine
        return o

    def _workspace(self) -> torch.Tensor:
        lib = _lib.lib()
        need = lib.pgp_workspace_size(self.N, self.D, self._prec)
        if self._ws is None or self._ws.numel() < need:
            self._ws = _lib.workspace(need, self.device)
        return self._ws

    def _grad_workspace(self) -> torch.Tensor:
        lib = _lib.lib()
        need = lib.pgp_grad_workspace_size(self.N, self.D)
        if self._gws is None or self._gws.numel() < need:
            self._gws = _lib.workspace(need, self.device)
        return self._gws

    def _stream(self):
        return _lib.stream_handle(self.device)

    def _vec(self):
        return torch.empty(self.N, dtype=torch.float64, device=self.device)

    def _check_status(self, res, what):
        """strict=True: raise on a fit or state whose numbers are not valid."""
        if not self.strict:
            return
        msg = numerical_failure(res.status, res.objective)
        if msg:
            raise FloatingPointError(f"{what}: {msg} (status {res.status})")

    def _check_bits(self, bits, what):
        """Status bits a predict or gradient call returned: strict=True raises,
        otherwise the outputs are NaN (the reference's behaviour)."""
        self.last_status = bits
        if bits and self.strict:
            raise FloatingPointError(f"{what}: {numerical_failure(bits)} (status {bits})")

    # ---- resident-state bookkeeping -----------------------------------------
    def _remember(self, key, weight, precision, result, fixed_point=False):
        """Record what the workspace holds.  Every state gets a new generation;
        `fixed_point` marks one formed at the fixed point of `key` (a fit, or
        a batched fit's weight), where the adjoint is valid."""
        self._generation = getattr(self, "_generation", 0) + 1
        self._state = dict(key=key, weight=weight, precision=precision,
                           wv=(weight.data_ptr(), weight._version),
                           pv=(precision.data_ptr(), precision._version),
                           gen=self._generation, fixed_point=bool(fixed_point))
        self.last_result = result

    def _state_matches(self, key, weight, precision) -> bool:
        st = self._state
        if st is None or st["key"] != key:
            return False
        if not (isinstance(weight, torch.Tensor) and isinstance(precision, torch.Tensor)):
            return False
        return ((weight.data_ptr(), weight._version) == st["wv"]
                and (precision.data_ptr(), precision._version) == st["pv"])

    def _as_device_vec(self, v) -> torch.Tensor:
        t = v if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v))
        return t.detach().to(torch.float64).reshape(-1).to(self.device).contiguous()

    def _set_state(self, ks, ls, weight, precision, fixed_point=False):
        """Factor B at a caller-given (weight, precision) (predict's input)."""
        lib = _lib.lib()
        ws = self._workspace()
        w = self._as_device_vec(weight)
        W = None if precision is None else self._as_device_vec(precision)
        Wout = self._vec() if W is None else None
        fout = self._vec()
        res = _lib.FitResultT()
        rc = lib.pgp_laplace_set_state(self._problem(None, Wout, fout), ks, ls, self._prec,
                                       w.data_ptr(), _lib.ptr(W), res, ws.data_ptr(), ws.numel(),
                                       self._stream())
        _lib.check(rc, "pgp_laplace_set_state")
        Wt = W if W is not None else Wout
        self._remember(self._key(ks, ls), w, Wt, res, fixed_point)
        self._check_status(res, "pgp_laplace_set_state")
        return w, Wt, fout, res

    # ---- public API ---------------------------------------------------------
    def _predict_setup(self, X_test, parameters, weight, precision):
        """Make the resident factor match (parameters, weight, precision) and
        stage X_test (M x D FP64 on the device) and the predict workspace."""
        lib = _lib.lib()
        ks, ls = self._structs(parameters)
        key = self._key(ks, ls)
        if not self._state_matches(key, weight, precision):
            _, Wt, _, res = self._set_state(ks, ls, weight, precision)
            if res.status & _lib.INFO_NEGATIVE_W:
                warnings.warn("precision has negative entries: the B = I + W^1/2 K W^1/2 form "
                              "is undefined there; predictive variances are not valid")
        Xt = torch.as_tensor(np.asarray(X_test) if not isinstance(X_test, torch.Tensor)
                             else X_test).detach().to(torch.float64)
        if Xt.dim() == 1:
            Xt = Xt[:, None]
        Xt = Xt.to(self.device).contiguous()
        M = Xt.shape[0]
        if Xt.shape[1] != self.D:
            raise ValueError(f"X_test has D={Xt.shape[1]}, expected {self.D}")
        need = lib.pgp_predict_workspace_size(self.N, M, self._prec)
        if self._pws is None or self._pws.numel() < need:
            self._pws = _lib.workspace(need, self.device)
        return lib, ks, Xt, M

    def predict(self, X_test, parameters, weight, precision):
        """Latent predictive mean and variance (probit/approximators.py:154-180)."""
        lib, ks, Xt, M = self._predict_setup(X_test, parameters, weight, precision)
        mean = torch.empty(M, dtype=torch.float64, device=self.device)
        var = torch.empty(M, dtype=torch.float64, device=self.device)
        ws = self._workspace()
        rc = lib.pgp_predict(ks, Xt.data_ptr(), M, mean.data_ptr(), var.data_ptr(),
                             ws.data_ptr(), ws.numel(), self._pws.data_ptr(), self._pws.numel(),
                             self._stream())
        bits = _lib.info_bits(rc, "pgp_predict")
        if bits:
            mean.fill_(float("nan"))
            var.fill_(float("nan"))
        self._check_bits(bits, "pgp_predict")
        return mean, var

    def predict_covariance(self, X_test, parameters, weight, precision):
        """Predictive covariance Kss - Kfs^T (K + W^-1)^-1 Kfs, an (M, M) tensor
        (probit/approximators.py:182-197)."""
        lib, ks, Xt, M = self._predict_setup(X_test, parameters, weight, precision)
        cov = torch.empty((M, M), dtype=torch.float64, device=self.device)
        ws = self._workspace()
        # column-major with ld = M is the transpose of torch's row-major
        # layout; the result is symmetric, so the same memory serves
        rc = lib.pgp_predict_covariance(ks, Xt.data_ptr(), M, cov.data_ptr(), M, ws.data_ptr(),
                                        ws.numel(), self._pws.data_ptr(), self._pws.numel(),
                                        self._stream())
        bits = _lib.info_bits(rc, "pgp_predict_covariance")
        if bits:
            cov.fill_(float("nan"))
        self._check_bits(bits, "pgp_predict_covariance")
        return cov

    def posterior_mean(self, weight):
        """The reference's posterior_mean (probit/approximators.py:199-202) reads an
        undefined `parameters` and cannot run; use precision(weight, parameters)."""
        raise NotImplementedError("posterior_mean needs parameters: use precision()")

    def approximate_posterior(self, parameters):
        """(weight, precision) of the approximate posterior (probit/approximators.py:204-210)."""
        return self._fit(parameters)

    def value_and_grad(self):
        """Value and gradient of the objective at the fixed point
        (probit/approximators.py:132-134).  Returns fn(parameters) ->
        (value, grads) with grads shaped like parameters (float64 NumPy)."""
        obj = self.objective()

        def fn(parameters):
            leaves, spec = _flatten(parameters)
            tl = [torch.as_tensor(np.asarray(v, dtype=np.float64) if not isinstance(v, torch.Tensor)
                                  else v.detach().cpu().to(torch.float64)).clone().requires_grad_(True)
                  for v in leaves]
            val = obj(_unflatten(spec, tl))
            grads = torch.autograd.grad(val, tl, allow_unused=True)
            out = [np.zeros(t.shape) if g is None else g.detach().cpu().numpy()
                   for t, g in zip(tl, grads)]
            return float(val.detach()), _unflatten(spec, out)

        return fn

    @abstractmethod
    def _fit(self, parameters):
        """Run the fixed point; return (weight, precision)."""

    @abstractmethod
    def construct(self):
        """The fixed-point map (parameters, weight) -> weight'."""

    @abstractmethod
    def objective(self):
        """The objective as fn(parameters) -> scalar tensor."""

    @abstractmethod
    def weight(self, parameters):
        """The weight a (f = K a)."""

    @abstractmethod
    def precision(self, weight, parameters):
        """(precision, posterior_mean)."""


class _LaplaceEvidence(torch.autograd.Function):
    """The objective at the fixed point (objective_LA or objective_VB) with
    its hand-written adjoint (gp._grad: pgp_laplace_grad / pgp_vb_grad)."""

    @staticmethod
    def forward(ctx, gp, parameters, var, ls, per, noise, cut):
        ks, lsd = gp._structs(parameters)
        pre = getattr(gp, "_prefit", {}).get(gp._key(ks, lsd))
        if pre is not None:  # fixed point from a batched fit: re-form its state
            res = gp._set_state(ks, lsd, pre, None, fixed_point=True)[3]
        else:
            res = gp._run_fit(ks, lsd)[2]
        ctx.gp = gp
        ctx.gen = gp._state["gen"]  # the fixed-point state the adjoint must use
        ctx.key = gp._key(ks, lsd)
        ctx.structs = (ks, lsd)
        ctx.parameters = parameters
        ctx.D = gp.D
        ctx.has_per = per is not None
        ctx.has_cut = cut is not None
        ctx.devs = [t.device if t is not None else None for t in (var, ls, per, noise, cut)]
        return torch.tensor(res.objective, dtype=torch.float64, device=var.device)

    @staticmethod
    def backward(ctx, g):
        gp = ctx.gp
        ks, lsd = ctx.structs
        st = gp._state
        # anything that re-formed the workspace since forward (predict,
        # precision or construct with other values, another fit) leaves a state
        # that is not this fixed point: fit again before the adjoint
        if st is None or st["gen"] != ctx.gen or st["key"] != ctx.key or not st["fixed_point"]:
            gp._run_fit(ks, lsd)
        vec = gp._grad(ks, lsd)
        D = ctx.D
        g = g.detach().cpu().double()
        o = 0
        gvar = torch.tensor(vec[o]) * g
        o += 1
        gls = torch.tensor(vec[o:o + D]) * g
        o += D
        gper = torch.tensor(vec[o:o + D]) * g if ctx.has_per else None
        o += D
        gnoise = torch.tensor(vec[o]) * g
        o += 1
        gcut = None
        if ctx.has_cut:
            J = lsd.J
            gcut = torch.tensor(vec[o:o + J + 1]) * g
        outs = [gvar, gls, gper, gnoise, gcut]
        outs = [None if t is None else t.to(d) for t, d in zip(outs, ctx.devs)]
        return (None, None, *outs)


class LaplaceGP(Approximator):
    """Laplace approximation (probit/approximators.py:213-277)."""

    def __repr__(self):
        return "LaplaceGP"

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self._prefit = {}

    def value_and_grad_batch(self, memory_fraction: float = 0.25):
        """``vmap(value_and_grad(objective))`` over a list of parameter sets
        (the lengthscale sweep of examples/classification.py:489-504).  The
        Newton fits of all sets run together through the batcher (one launch
        per blocked step for every set, the serving path's kernels); then, set
        by set, B is re-formed and factored at its fixed point and the adjoint
        gives the gradient.  Returns fn(list_of_parameters) -> (values (K,)
        float64 NumPy, [grads shaped like each parameters]).  At most
        ``memory_fraction`` of the free device memory holds fit workspaces at
        once (larger sweeps run in chunks)."""
        single = self.value_and_grad()

        def fn(parameter_list):
            plist = list(parameter_list)
            structs = [self._structs(p) for p in plist]
            weights = self._batched_fits(structs, memory_fraction)
            values, grads = [], []
            try:
                for p, (ks, ls), w in zip(plist, structs, weights):
                    self._prefit = {self._key(ks, ls): w}
                    v, g = single(p)
                    values.append(v)
                    grads.append(g)
            finally:
                self._prefit = {}
            return np.asarray(values, dtype=np.float64), grads

        return fn

    def _batched_fits(self, structs, memory_fraction):
        """Weights at the fixed points of `structs` ((kernel, likelihood)
        pairs) from the batcher, in order."""
        lib = _lib.lib()
        need = int(lib.pgp_workspace_size(self.N, self.D, self._prec))
        free, _ = torch.cuda.mem_get_info(self.device)
        chunk = max(1, min(len(structs), int(free * memory_fraction) // max(1, need), 512))
        out = []
        for c0 in range(0, len(structs), chunk):
            part = structs[c0:c0 + chunk]
            n = len(part)
            tb = lib.pgp_batcher_table_size(n)
            table = _lib.workspace(tb, self.device)
            h = C.c_void_p()
            _lib.check(lib.pgp_batcher_create(n, self._options(), table.data_ptr(), tb,
                                              self._stream(), C.byref(h)), "pgp_batcher_create")
            try:
                items, keep = [], []
                for i, (ks, ls) in enumerate(part):
                    ws = _lib.workspace(need, self.device)
                    w, prec = self._vec(), self._vec()
                    item = _lib.BatchItemT()
                    item.problem = self._problem(w, prec)
                    item.kernel = ks
                    item.lik = ls
                    item.workspace = ws.data_ptr()
                    item.workspace_bytes = ws.numel()
                    _lib.check(lib.pgp_batcher_submit(h, item, i), "pgp_batcher_submit")
                    items.append(item)
                    keep.append((ws, w, prec))
                done = (C.c_int64 * n)()
                nd = C.c_int32(0)
                finished = 0
                while finished < n:
                    _lib.check(lib.pgp_batcher_step(h, done, n, C.byref(nd)), "pgp_batcher_step")
                    finished += nd.value
                for it in items:
                    self._check_status(it.result, "batched pgp_laplace_fit")
                    if it.result.status & _lib.INFO_MAXITER:
                        warnings.warn(f"Newton iteration stopped at maxiter={self.maxiter}")
                out.extend(k[1] for k in keep)
            finally:
                lib.pgp_batcher_destroy(h)
        return out

    def _run_fit(self, ks, ls):
        lib = _lib.lib()
        ws = self._workspace()
        weight, precision = self._vec(), self._vec()
        res = _lib.FitResultT()
        rc = lib.pgp_laplace_fit(self._problem(weight, precision), ks, ls, self._options(), res,
                                 ws.data_ptr(), ws.numel(), self._stream())
        _lib.check(rc, "pgp_laplace_fit")
        self._remember(self._key(ks, ls), weight, precision, res, fixed_point=True)
        self._check_status(res, "pgp_laplace_fit")
        if res.status & _lib.INFO_MAXITER:
            warnings.warn(f"Newton iteration stopped at maxiter={self.maxiter} "
                          f"(step {res.step_norm:.3e} > tol {self.tolerance:.1e})")
        return weight, precision, res

    def _fit(self, parameters):
        ks, ls = self._structs(parameters)
        w, p, _ = self._run_fit(ks, ls)
        return w, p

    def _grad(self, ks, ls):
        _refuse_closed_form_gradient(self.lik_kind)
        lib = _lib.lib()
        ws = self._workspace()
        gws = self._grad_workspace()
        out = (np.zeros(2 + 2 * self.D + _lib.MAX_CUT)).astype(np.float64)
        rc = lib.pgp_laplace_grad(ks, ls, out.ctypes.data, ws.data_ptr(), ws.numel(),
                                  gws.data_ptr(), gws.numel(), self._stream())
        self._check_bits(_lib.info_bits(rc, "pgp_laplace_grad"), "pgp_laplace_grad")
        return out

    def construct(self):
        """f_LA: (parameters, weight) -> grad log p(y | K weight)
        (probit/approximators.py:238-246, probit/implicit/Laplace.py:4-9)."""

        def f(parameters, weight):
            ks, ls = self._structs(parameters)
            _, _, latent, _ = self._set_state(ks, ls, weight, None)
            d1 = torch.empty_like(latent)
            rc = _lib.lib().pgp_likelihood_eval(ls, latent.data_ptr(), self._y.data_ptr(), self.N,
                                                None, d1.data_ptr(), None, self._stream())
            _lib.check(rc, "pgp_likelihood_eval")
            return d1

        return f

    def objective(self):
        """-log q(y | X, theta) at the Laplace fixed point
        (probit/approximators.py:248-263, probit/implicit/Laplace.py:12-30),
        differentiable through torch autograd via the hand-written adjoint."""

        def obj(parameters):
            kern = self._kernel(parameters[0])
            var, ls, per = kern.spec_tensors(self.D)
            lp = parameters[1]
            lp = lp if isinstance(lp, (tuple, list)) else (lp,)
            noise = torch.as_tensor(lp[0], dtype=torch.float64) if not isinstance(
                lp[0], torch.Tensor) else lp[0].to(torch.float64)
            cut = None
            if self.lik_kind != _lib.LIK_GAUSSIAN:
                c = lp[1]
                cut = c.to(torch.float64) if isinstance(c, torch.Tensor) else torch.as_tensor(
                    np.asarray(c, dtype=np.float64))
            return _LaplaceEvidence.apply(self, parameters, var, ls, per, noise.reshape(()),
                                          cut)

        return obj

    def weight(self, parameters):
        """Newton fixed point from zeros (probit/approximators.py:265-269)."""
        return self._fit(parameters)[0]

    def precision(self, weight, parameters):
        """(W, f) with f = K weight, W = -d2 log p (probit/approximators.py:271-277)."""
        ks, ls = self._structs(parameters)
        _, W, f, _ = self._set_state(ks, ls, weight, None)
        return W, f


class VBGP(Approximator):
    """Variational Bayes approximation (probit/approximators.py:280-339)."""

    def __repr__(self):
        return "VBGP"

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)

    def _run_fit(self, ks, ls):
        lib = _lib.lib()
        ws = self._workspace()
        gws = self._grad_workspace()
        weight, precision = self._vec(), self._vec()
        res = _lib.FitResultT()
        rc = lib.pgp_vb_fit(self._problem(weight, precision), ks, ls, self._options(), res,
                            ws.data_ptr(), ws.numel(), gws.data_ptr(), gws.numel(),
                            self._stream())
        _lib.check(rc, "pgp_vb_fit")
        self._remember(self._key(ks, ls), weight, precision, res, fixed_point=True)
        self._check_status(res, "pgp_vb_fit")
        if res.status & _lib.INFO_MAXITER:
            warnings.warn(f"fixed-point iteration stopped at maxiter={self.maxiter}")
        return weight, precision, res

    def _fit(self, parameters):
        ks, ls = self._structs(parameters)
        w, p, _ = self._run_fit(ks, ls)
        return w, p

    def construct(self):
        """f_VB: (parameters, weight) -> (sigma^2 I + K)^-1 (K weight + sigma
        d log p(y | K weight)) (probit/approximators.py:306-314,
        probit/implicit/VB.py:4-16)."""

        def f(parameters, weight):
            ks, ls = self._structs(parameters)
            lib = _lib.lib()
            ws = self._workspace()
            w = self._as_device_vec(weight)
            out = self._vec()
            rc = lib.pgp_vb_map(self._problem(out), ks, ls, self._prec, w.data_ptr(),
                                ws.data_ptr(), ws.numel(), self._stream())
            _lib.check(rc, "pgp_vb_map")
            self._state = None  # the workspace now holds sigma^2 I + K
            return out

        return f

    def _grad(self, ks, ls):
        _refuse_closed_form_gradient(self.lik_kind)
        lib = _lib.lib()
        ws = self._workspace()
        gws = self._grad_workspace()
        out = (np.zeros(2 + 2 * self.D + _lib.MAX_CUT)).astype(np.float64)
        rc = lib.pgp_vb_grad(ks, ls, out.ctypes.data, ws.data_ptr(), ws.numel(),
                             gws.data_ptr(), gws.numel(), self._stream())
        self._check_bits(_lib.info_bits(rc, "pgp_vb_grad"), "pgp_vb_grad")
        return out

    def objective(self):
        """objective_VB (probit/implicit/VB.py:19-40) at the VB fixed point,
        differentiable through torch autograd via the hand-written adjoint
        (pgp_vb_grad: the implicit-function gradient fixed_point_layer gives,
        probit/implicit/solvers.py:52-64)."""

        def obj(parameters):
            kern = self._kernel(parameters[0])
            var, ls, per = kern.spec_tensors(self.D)
            lp = parameters[1]
            lp = lp if isinstance(lp, (tuple, list)) else (lp,)
            noise = torch.as_tensor(lp[0], dtype=torch.float64) if not isinstance(
                lp[0], torch.Tensor) else lp[0].to(torch.float64)
            cut = None
            if self.lik_kind != _lib.LIK_GAUSSIAN:
                c = lp[1]
                cut = c.to(torch.float64) if isinstance(c, torch.Tensor) else torch.as_tensor(
                    np.asarray(c, dtype=np.float64))
            return _LaplaceEvidence.apply(self, parameters, var, ls, per, noise.reshape(()),
                                          cut)

        return obj

    def weight(self, parameters):
        return self._fit(parameters)[0]

    def precision(self, weight, parameters):
        """(1/sigma^2, K weight) (probit/approximators.py:336-339)."""
        ks, ls = self._structs(parameters)
        w = self._as_device_vec(weight)
        _, _, f, _ = self._set_state(ks, ls, w, torch.full((self.N,), 1.0 / ls.noise_std**2,
                                                           dtype=torch.float64,
                                                           device=self.device))
        return torch.full((self.N,), 1.0 / ls.noise_std**2, dtype=torch.float64,
                          device=self.device), f
