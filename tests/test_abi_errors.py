"""Argument checking of every C-ABI entry point: each invalid call returns
its negative status and sets pgp_last_error() before any HIP call, so these
run without a GPU.  The Python layer turns such a status into an exception
(probit_amd._lib.check); the numeric failures of a valid call (NaN, status
bits) are covered by the GPU tests."""

from __future__ import annotations

import ctypes as C
import math

import pytest

FAKE = 0x1000  # a non-NULL pointer that no failing call dereferences


def _lib():
    from probit_amd import _lib
    return _lib, _lib.load()


def _kernel(L, dim=2, kind=None, variance=1.0, ls=1.0, period=0.0):
    k = L.KernelT()
    k.kind = L.KERNEL_EQ if kind is None else kind
    k.dim = dim
    k.variance = variance
    for d in range(L.MAX_DIM):
        k.lengthscale[d] = ls
        k.period[d] = period
    return k


def _gauss(L, noise=0.3):
    lk = L.LikelihoodT()
    lk.kind = L.LIK_GAUSSIAN
    lk.noise_std = noise
    return lk


def _probit(L, cut=(-math.inf, 0.0, math.inf), noise=0.5):
    lk = L.LikelihoodT()
    lk.kind = L.LIK_PROBIT
    lk.J = len(cut) - 1
    lk.noise_std = noise
    for j, c in enumerate(cut):
        lk.cutpoints[j] = c
    return lk


def _problem(L, N=100, D=2, X=FAKE, y=FAKE, weight=FAKE):
    p = L.ProblemT()
    p.N, p.D, p.X, p.y, p.weight = N, D, X, y, weight
    return p


def _opts(L, precision=0):
    o = L.FitOptionsT()
    o.tolerance, o.maxiter, o.precision, o.refine = 1e-5, 100, precision, -1
    return o


def _fit(L, lib, problem=None, kernel=None, lik=None, opts=None, result=None, ws=None, nbytes=0):
    r = L.FitResultT()
    return lib.pgp_laplace_fit(problem, kernel, lik, opts, result if result is not None else r,
                               ws, nbytes, None)


def _err(lib) -> str:
    return lib.pgp_last_error().decode()


# (name, call(L, lib) -> rc, expected rc, substring of the error message)
CASES = [
    ("fit: problem NULL", lambda L, lib: _fit(L, lib, None, _kernel(L), _gauss(L), _opts(L)), -1,
     "problem is NULL"),
    ("fit: N = 0", lambda L, lib: _fit(L, lib, _problem(L, N=0), _kernel(L), _gauss(L), _opts(L)),
     -1, "N must be positive"),
    ("fit: D too large", lambda L, lib: _fit(L, lib, _problem(L, D=L.MAX_DIM + 1),
                                             _kernel(L, dim=L.MAX_DIM + 1), _gauss(L), _opts(L)),
     -1, "D must be in"),
    ("fit: X NULL", lambda L, lib: _fit(L, lib, _problem(L, X=None), _kernel(L), _gauss(L),
                                        _opts(L)), -1, "device pointers"),
    ("fit: kernel NULL", lambda L, lib: _fit(L, lib, _problem(L), None, _gauss(L), _opts(L)), -2,
     "kernel is NULL"),
    ("fit: kernel kind", lambda L, lib: _fit(L, lib, _problem(L), _kernel(L, kind=7), _gauss(L),
                                             _opts(L)), -2, "unknown kernel kind"),
    ("fit: kernel dim", lambda L, lib: _fit(L, lib, _problem(L), _kernel(L, dim=3), _gauss(L),
                                            _opts(L)), -2, "!= problem D"),
    ("fit: variance 0", lambda L, lib: _fit(L, lib, _problem(L), _kernel(L, variance=0.0),
                                            _gauss(L), _opts(L)), -2, "variance must be > 0"),
    ("fit: lengthscale NaN", lambda L, lib: _fit(L, lib, _problem(L), _kernel(L, ls=math.nan),
                                                 _gauss(L), _opts(L)), -2, "lengthscale"),
    ("fit: period < 0", lambda L, lib: _fit(L, lib, _problem(L), _kernel(L, period=-1.0),
                                            _gauss(L), _opts(L)), -2, "period"),
    ("fit: likelihood NULL", lambda L, lib: _fit(L, lib, _problem(L), _kernel(L), None, _opts(L)),
     -3, "likelihood is NULL"),
    ("fit: noise 0", lambda L, lib: _fit(L, lib, _problem(L), _kernel(L), _gauss(L, 0.0),
                                         _opts(L)), -3, "noise_std"),
    ("fit: one class", lambda L, lib: _fit(L, lib, _problem(L), _kernel(L),
                                           _probit(L, (-math.inf, math.inf)), _opts(L)), -3,
     "J must be in"),
    ("fit: first cutpoint finite", lambda L, lib: _fit(L, lib, _problem(L), _kernel(L),
                                                       _probit(L, (-5.0, 0.0, math.inf)),
                                                       _opts(L)), -3, "-inf"),
    ("fit: cutpoints descending", lambda L, lib: _fit(L, lib, _problem(L), _kernel(L),
                                                      _probit(L, (-math.inf, 1.0, -1.0,
                                                                  math.inf)), _opts(L)), -3,
     "ascending"),
    ("fit: options NULL", lambda L, lib: _fit(L, lib, _problem(L), _kernel(L), _gauss(L), None),
     -4, "options are NULL"),
    ("fit: precision", lambda L, lib: _fit(L, lib, _problem(L), _kernel(L), _gauss(L),
                                           _opts(L, precision=3)), -4, "unknown precision"),
    ("fit: workspace NULL", lambda L, lib: _fit(L, lib, _problem(L), _kernel(L), _gauss(L),
                                                _opts(L)), -6, "workspace is NULL"),
    ("fit: workspace small", lambda L, lib: _fit(L, lib, _problem(L), _kernel(L), _gauss(L),
                                                 _opts(L), ws=FAKE, nbytes=64), -7,
     "workspace too small"),
    ("vb_fit: options NULL", lambda L, lib: lib.pgp_vb_fit(_problem(L), _kernel(L), _gauss(L),
                                                           None, L.FitResultT(), None, 0, None,
                                                           0, None), -4, "NULL"),
    ("vb_fit: workspace small", lambda L, lib: lib.pgp_vb_fit(
        _problem(L), _kernel(L), _gauss(L), _opts(L), L.FitResultT(), FAKE, 64, None, 0, None),
     -7, "workspace too small"),
    ("vb_map: weight NULL", lambda L, lib: lib.pgp_vb_map(_problem(L), _kernel(L), _gauss(L), 0,
                                                          None, None, 0, None), -5,
     "weight is NULL"),
    ("vb_map: output NULL", lambda L, lib: lib.pgp_vb_map(_problem(L, weight=None), _kernel(L),
                                                          _gauss(L), 0, FAKE, None, 0, None),
     -1, "problem->weight"),
    ("vb_map: workspace small", lambda L, lib: lib.pgp_vb_map(_problem(L), _kernel(L), _gauss(L),
                                                              0, FAKE, FAKE, 64, None), -6,
     "workspace too small"),
    ("set_state: weight NULL", lambda L, lib: lib.pgp_laplace_set_state(
        _problem(L), _kernel(L), _gauss(L), 0, None, None, L.FitResultT(), None, 0, None), -5,
     "weight is NULL"),
    ("set_state: workspace small", lambda L, lib: lib.pgp_laplace_set_state(
        _problem(L), _kernel(L), _gauss(L), 0, FAKE, FAKE, L.FitResultT(), FAKE, 64, None), -9,
     "workspace too small"),
    ("predict: no test points", lambda L, lib: lib.pgp_predict(_kernel(L), FAKE, 0, FAKE, FAKE,
                                                               FAKE, 0, FAKE, 0, None), -2,
     "Xtest / M"),
    ("predict: outputs NULL", lambda L, lib: lib.pgp_predict(_kernel(L), FAKE, 10, None, FAKE,
                                                             FAKE, 0, FAKE, 0, None), -4,
     "mean / var"),
    ("predict: workspace NULL", lambda L, lib: lib.pgp_predict(_kernel(L), FAKE, 10, FAKE, FAKE,
                                                               None, 0, FAKE, 0, None), -6,
     "workspace is NULL"),
    ("covariance: ld < M", lambda L, lib: lib.pgp_predict_covariance(
        _kernel(L), FAKE, 10, FAKE, 9, FAKE, 0, FAKE, 0, None), -5, "ld < M"),
    ("covariance: cov NULL", lambda L, lib: lib.pgp_predict_covariance(
        _kernel(L), FAKE, 10, None, 10, FAKE, 0, FAKE, 0, None), -4, "cov is NULL"),
    ("laplace_grad: output NULL", lambda L, lib: lib.pgp_laplace_grad(
        _kernel(L), _gauss(L), None, FAKE, 0, FAKE, 0, None), -3, "grad_out is NULL"),
    ("laplace_grad: workspace NULL", lambda L, lib: lib.pgp_laplace_grad(
        _kernel(L), _gauss(L), FAKE, None, 0, FAKE, 0, None), -4, "workspace is NULL"),
    ("vb_grad: output NULL", lambda L, lib: lib.pgp_vb_grad(
        _kernel(L), _gauss(L), None, FAKE, 0, FAKE, 0, None), -3, "grad_out is NULL"),
    ("vb_grad: workspace NULL", lambda L, lib: lib.pgp_vb_grad(
        _kernel(L), _gauss(L), FAKE, None, 0, FAKE, 0, None), -4, "workspace is NULL"),
    ("predict: kernel NULL", lambda L, lib: lib.pgp_predict(None, FAKE, 10, FAKE, FAKE, FAKE, 0,
                                                            FAKE, 0, None), -1,
     "kernel is NULL"),
    ("covariance: kernel NULL", lambda L, lib: lib.pgp_predict_covariance(
        None, FAKE, 10, FAKE, 10, FAKE, 0, FAKE, 0, None), -1, "kernel is NULL"),
    ("laplace_grad: kernel NULL", lambda L, lib: lib.pgp_laplace_grad(
        None, _gauss(L), FAKE, FAKE, 0, FAKE, 0, None), -1, "kernel is NULL"),
    ("laplace_grad: likelihood NULL", lambda L, lib: lib.pgp_laplace_grad(
        _kernel(L), None, FAKE, FAKE, 0, FAKE, 0, None), -2, "likelihood is NULL"),
    ("vb_grad: kernel NULL", lambda L, lib: lib.pgp_vb_grad(
        None, _gauss(L), FAKE, FAKE, 0, FAKE, 0, None), -1, "kernel is NULL"),
    ("vb_grad: likelihood NULL", lambda L, lib: lib.pgp_vb_grad(
        _kernel(L), None, FAKE, FAKE, 0, FAKE, 0, None), -2, "likelihood is NULL"),
    ("noalloc: N beyond the limit", lambda L, lib: lib.pgp_laplace_fit_noalloc(
        _problem(L, N=L.NOALLOC_MAX_NP + 1), _kernel(L), _gauss(L), _opts(L), L.FitResultT(),
        None), -1, "no-allocation limit"),
    ("noalloc: problem NULL", lambda L, lib: lib.pgp_laplace_fit_noalloc(
        None, _kernel(L), _gauss(L), _opts(L), L.FitResultT(), None), -1, "problem is NULL"),
    ("noalloc: options NULL", lambda L, lib: lib.pgp_laplace_fit_noalloc(
        _problem(L), _kernel(L), _gauss(L), None, L.FitResultT(), None), -4,
     "options are NULL"),
    ("metrics: P NULL", lambda L, lib: lib.pgp_ordinal_metrics(None, 3, FAKE, 10, FAKE, None,
                                                               FAKE, FAKE, None), -1,
     "P is NULL"),
    ("metrics: J", lambda L, lib: lib.pgp_ordinal_metrics(FAKE, 1, FAKE, 10, FAKE, None, FAKE,
                                                          FAKE, None), -2, "J must be"),
    ("metrics: workspace NULL", lambda L, lib: lib.pgp_ordinal_metrics(
        FAKE, 3, FAKE, 10, FAKE, None, FAKE, None, None), -8, "workspace"),
    ("tune: precision", lambda L, lib: lib.pgp_tune_get(3, 0, L.TuneEntryT()), -1,
     "unknown precision"),
    ("tune: class", lambda L, lib: lib.pgp_tune_get(1, 2, L.TuneEntryT()), -2, "size class"),
    ("tune: superstep 0", lambda L, lib: lib.pgp_tune_set(1, 0, L.TuneEntryT(0, 512, 512, 0)),
     -3, "superstep"),
    ("tune: superstep 17", lambda L, lib: lib.pgp_tune_set(1, 0, L.TuneEntryT(17, 512, 512, 0)),
     -3, "superstep"),
    ("tune: negative column_max", lambda L, lib: lib.pgp_tune_set(
        1, 1, L.TuneEntryT(8, -1, 512, 0)), -3, "column_max"),
    ("status mapping: result NULL", lambda L, lib: lib.pgp_fit_result_apply_info(None, 1), -1,
     "result is NULL"),
    ("gram: kernel NULL", lambda L, lib: lib.pgp_gram(None, 0, FAKE, 10, FAKE, 10, FAKE, 10, None),
     -1, "kernel is NULL"),
    ("gram: dtype", lambda L, lib: lib.pgp_gram(_kernel(L), 2, FAKE, 10, FAKE, 10, FAKE, 10, None),
     -2, "dtype"),
    ("gram: N = 0", lambda L, lib: lib.pgp_gram(_kernel(L), 0, FAKE, 0, FAKE, 10, FAKE, 10, None),
     -3, "X / N"),
    ("gram: M = 0", lambda L, lib: lib.pgp_gram(_kernel(L), 0, FAKE, 10, FAKE, 0, FAKE, 10, None),
     -5, "Z / M"),
    ("gram: ld < N", lambda L, lib: lib.pgp_gram(_kernel(L), 0, FAKE, 10, FAKE, 10, FAKE, 9, None),
     -8, "ld"),
    ("gram: dim 0", lambda L, lib: lib.pgp_gram(_kernel(L, dim=0), 0, FAKE, 10, FAKE, 10, FAKE, 10,
                                                None), -1, "dim out of range"),
    ("likelihood: NULL", lambda L, lib: lib.pgp_likelihood_eval(None, FAKE, FAKE, 10, FAKE, FAKE,
                                                                FAKE, None), -1,
     "likelihood is NULL"),
    ("likelihood: N < 0", lambda L, lib: lib.pgp_likelihood_eval(_gauss(L), FAKE, FAKE, -1, FAKE,
                                                                 FAKE, FAKE, None), -4, "N < 0"),
    ("probit predictive: Gaussian", lambda L, lib: lib.pgp_probit_predictive(
        _gauss(L), FAKE, FAKE, 10, FAKE, None), -1, "probit likelihood"),
    ("probit predictive: NULL", lambda L, lib: lib.pgp_probit_predictive(
        _probit(L), None, FAKE, 10, FAKE, None), -2, "NULL pointer"),
    ("potrf: N not a multiple of 64", lambda L, lib: lib.pgp_potrf(1, FAKE, 100, 128, FAKE, None,
                                                                   FAKE, None), -3,
     "multiple of 64"),
    ("potrf: info NULL", lambda L, lib: lib.pgp_potrf(1, FAKE, 128, 128, FAKE, None, None, None),
     -7, "info"),
    ("trsv: precision", lambda L, lib: lib.pgp_trsv(3, FAKE, 64, 64, FAKE, FAKE, FAKE, 0, None),
     -1, "unknown precision"),
    ("trsv: ld < N", lambda L, lib: lib.pgp_trsv(0, FAKE, 128, 64, FAKE, FAKE, FAKE, 0, None), -4,
     "ld < N"),
    ("trsm: M not a multiple of 64", lambda L, lib: lib.pgp_trsm(0, FAKE, 64, 64, FAKE, FAKE, 10,
                                                                 64, None), -7,
     "M must be a positive multiple of 64"),
    ("prior: jitter < 0", lambda L, lib: lib.pgp_prior_sample(
        _kernel(L), FAKE, 10, -1.0, FAKE, 1, 10, FAKE, 10, FAKE, 1 << 30, None), -4, "jitter"),
    ("prior: jitter NaN", lambda L, lib: lib.pgp_prior_sample(
        _kernel(L), FAKE, 10, math.nan, FAKE, 1, 10, FAKE, 10, FAKE, 1 << 30, None), -4,
     "jitter"),
    ("prior: ldz < N", lambda L, lib: lib.pgp_prior_sample(
        _kernel(L), FAKE, 10, 0.0, FAKE, 1, 9, FAKE, 10, FAKE, 1 << 30, None), -5, "ldz"),
    ("prior: no samples", lambda L, lib: lib.pgp_prior_sample(
        _kernel(L), FAKE, 10, 0.0, FAKE, 0, 10, FAKE, 10, FAKE, 1 << 30, None), -5, "S"),
    ("prior: ldf < N", lambda L, lib: lib.pgp_prior_sample(
        _kernel(L), FAKE, 10, 0.0, FAKE, 1, 10, FAKE, 9, FAKE, 1 << 30, None), -8, "ldf"),
    ("prior: workspace small", lambda L, lib: lib.pgp_prior_sample(
        _kernel(L), FAKE, 10, 0.0, FAKE, 1, 10, FAKE, 10, FAKE, 64, None), -10,
     "workspace too small"),
    ("batcher: options NULL", lambda L, lib: lib.pgp_batcher_create(4, None, FAKE, 1 << 20, None,
                                                                    C.byref(C.c_void_p())), -2,
     "options are NULL"),
    ("batcher: table small", lambda L, lib: lib.pgp_batcher_create(4, _opts(L), FAKE, 8, None,
                                                                   C.byref(C.c_void_p())), -3,
     "table too small"),
    ("batcher: submit to NULL", lambda L, lib: lib.pgp_batcher_submit(None, L.BatchItemT(), 0), -1,
     "batcher is NULL"),
    ("batcher: step of NULL", lambda L, lib: lib.pgp_batcher_step(None, None, 0,
                                                                  C.byref(C.c_int32())), -1,
     "batcher is NULL"),
    ("batcher: finish without n_done", lambda L, lib: lib.pgp_batcher_finish(FAKE, None, 0, None),
     -4, "n_done is NULL"),
]


@pytest.mark.parametrize("name,call,code,msg", CASES, ids=[c[0] for c in CASES])
def test_invalid_call_is_rejected(name, call, code, msg):
    L, lib = _lib()
    rc = call(L, lib)
    assert rc == code, (name, rc, _err(lib))
    assert msg in _err(lib), (name, _err(lib))


def test_size_queries_reject_invalid_sizes():
    L, lib = _lib()
    assert lib.pgp_workspace_size(-5, 2, 0) == -1
    assert lib.pgp_workspace_size(100, 2, 3) == -1
    assert lib.pgp_predict_workspace_size(100, 0, 0) == -1
    assert lib.pgp_grad_workspace_size(0, 2) == -1
    assert lib.pgp_prior_sample_workspace_size(0) == -1
    # sizes grow with N and are padded to whole 64-row tiles
    assert lib.pgp_workspace_size(64, 2, 1) == lib.pgp_workspace_size(1, 2, 1)
    assert lib.pgp_workspace_size(65, 2, 1) > lib.pgp_workspace_size(64, 2, 1)
    assert lib.pgp_prior_sample_workspace_size(65) > lib.pgp_prior_sample_workspace_size(64)


def test_null_batcher_queries_are_harmless():
    L, lib = _lib()
    assert lib.pgp_batcher_in_flight(None) == 0
    assert lib.pgp_batcher_launched(None) == 0
    assert lib.pgp_batcher_ready(None) == 0
    assert lib.pgp_batcher_destroy(None) == 0


def test_python_layer_raises_on_a_rejected_call():
    L, lib = _lib()
    rc = lib.pgp_gram(None, 0, FAKE, 10, FAKE, 10, FAKE, 10, None)
    with pytest.raises(RuntimeError, match="kernel is NULL"):
        L.check(rc, "pgp_gram")


@pytest.mark.parametrize("word,prior,bits,nan", [
    (0, 0, 0, False),
    (1, 0, 1, True),       # non-positive pivot: NaN objective
    (8, 0, 8, True),       # an expired bounded wait: NaN objective
    (9, 0, 9, True),
    (16, 0, 0, False),     # other bits of the word are not status of the numbers
    (0, 2, 2, True),       # negative W found on the host: NaN objective
    (0, 4, 4, False),      # maxiter alone leaves the objective a number
    (0, 16, 16, False),    # BF16 escalation is informational
    (8, 16, 24, True),
])
def test_fit_status_mapping(word, prior, bits, nan):
    """The device status word of a fit becomes result.status bits and a NaN
    objective exactly where the numbers are invalid (host only, no GPU)."""
    L, lib = _lib()
    r = L.FitResultT()
    r.status = prior
    r.objective = 12.5
    assert lib.pgp_fit_result_apply_info(C.byref(r), word) == 0
    assert r.status == bits
    assert math.isnan(r.objective) == nan
    if not nan:
        assert r.objective == 12.5


def test_tune_table_round_trips_on_the_host():
    """The schedule table needs no GPU: set, get back, restore."""
    L, lib = _lib()
    old = L.TuneEntryT()
    assert lib.pgp_tune_get(2, 1, old) == 0
    assert 1 <= old.superstep <= 16
    assert lib.pgp_tune_set(2, 1, L.TuneEntryT(5, 100, 0, 0)) == 0
    got = L.TuneEntryT()
    assert lib.pgp_tune_get(2, 1, got) == 0
    assert (got.superstep, got.column_max, got.fuse_max, got.column_merge) == (5, 100, 0, 0)
    assert lib.pgp_tune_set(2, 1, L.TuneEntryT(5, 100, 0, 3)) == -3  # column_merge is 0, 1 or 2
    other = L.TuneEntryT()
    assert lib.pgp_tune_get(2, 0, other) == 0 and other.superstep == old.superstep
    assert lib.pgp_tune_set(2, 1, old) == 0
