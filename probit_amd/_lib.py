"""ctypes binding of libprobit_amd.so (the C ABI declared in include/probit_amd.h).

The shared library is the only compute path: if it is missing or no GPU is
visible, every entry point raises instead of falling back to anything else.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import torch

MAX_DIM = 32
MAX_CUT = 17
NOALLOC_MAX_NP = 1024

KERNEL_EQ = 0
KERNEL_MATERN12 = 1

LIK_GAUSSIAN = 0
LIK_PROBIT = 1
LIK_PROBIT_CLOSED = 2

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16 = 2
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16": PREC_BF16}

INFO_NOT_PD = 1
INFO_NEGATIVE_W = 2
INFO_MAXITER = 4
INFO_STALLED = 8
INFO_ESCALATED = 16

FIT_NO_GRAPH = 1

TIMER_SYRK = 0
TIMER_PANEL = 1
TIMER_TRSV = 2
TIMER_GRAM = 3
TIMER_TRSM = 4
TIMER_REFINE = 5
TIMER_SUPER = 6

LIB_PATH = Path(__file__).resolve().parent / "libprobit_amd.so"


class KernelT(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("dim", C.c_int32),
        ("reserved0", C.c_int32),
        ("reserved1", C.c_int32),
        ("variance", C.c_double),
        ("lengthscale", C.c_double * MAX_DIM),
        ("period", C.c_double * MAX_DIM),
    ]


class LikelihoodT(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("J", C.c_int32),
        ("single_precision_bounds", C.c_int32),
        ("reserved", C.c_int32),
        ("noise_std", C.c_double),
        ("cutpoints", C.c_double * MAX_CUT),
    ]


class FitOptionsT(C.Structure):
    _fields_ = [
        ("tolerance", C.c_double),
        ("maxiter", C.c_int32),
        ("precision", C.c_int32),
        ("refine", C.c_int32),
        ("flags", C.c_int32),
    ]


class FitResultT(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("status", C.c_int32),
        ("step_norm", C.c_double),
        ("sum_log_lik", C.c_double),
        ("half_f_dot_a", C.c_double),
        ("log_det", C.c_double),
        ("objective", C.c_double),
    ]


class ProblemT(C.Structure):
    _fields_ = [
        ("N", C.c_int64),
        ("D", C.c_int32),
        ("reserved", C.c_int32),
        ("X", C.c_void_p),
        ("y", C.c_void_p),
        ("weight", C.c_void_p),
        ("precision", C.c_void_p),
        ("latent", C.c_void_p),
    ]


class BatchItemT(C.Structure):
    _fields_ = [
        ("problem", ProblemT),
        ("kernel", KernelT),
        ("lik", LikelihoodT),
        ("workspace", C.c_void_p),
        ("workspace_bytes", C.c_int64),
        ("Xtest", C.c_void_p),
        ("M", C.c_int64),
        ("mean", C.c_void_p),
        ("var", C.c_void_p),
        ("predict_workspace", C.c_void_p),
        ("predict_workspace_bytes", C.c_int64),
        ("result", FitResultT),
    ]


TUNE_CLASSES = 2
TUNE_CLASS_NP = 1024


class TuneEntryT(C.Structure):
    _fields_ = [
        ("superstep", C.c_int32),
        ("column_max", C.c_int32),
        ("fuse_max", C.c_int32),
        ("column_merge", C.c_int32),
    ]


# name -> (restype, argtypes)
_P = C.c_void_p
_I64 = C.c_int64
_I32 = C.c_int32
_SIGNATURES = {
    "pgp_last_error": (C.c_char_p, []),
    "pgp_version": (C.c_int, []),
    "pgp_tune_get": (C.c_int, [_I32, _I32, C.POINTER(TuneEntryT)]),
    "pgp_tune_set": (C.c_int, [_I32, _I32, C.POINTER(TuneEntryT)]),
    "pgp_fit_result_apply_info": (C.c_int, [C.POINTER(FitResultT), C.c_int32]),
    "pgp_timer_enable": (C.c_int, [C.c_uint32]),
    "pgp_timer_reset": (C.c_int, []),
    "pgp_timer_read": (C.c_int, [C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "pgp_gram": (C.c_int, [C.POINTER(KernelT), C.c_int, _P, _I64, _P, _I64, _P, _I64, _P]),
    "pgp_likelihood_eval": (C.c_int, [C.POINTER(LikelihoodT), _P, _P, _I64, _P, _P, _P, _P]),
    "pgp_probit_predictive": (C.c_int, [C.POINTER(LikelihoodT), _P, _P, _I64, _P, _P]),
    "pgp_ordinal_metrics": (C.c_int, [_P, _I32, _P, _I64, _P, _P, _P, _P, _P]),
    "pgp_potrf": (C.c_int, [C.c_int, _P, _I64, _I64, _P, _P, _P, _P]),
    "pgp_prior_sample_workspace_size": (_I64, [_I64]),
    "pgp_prior_sample": (C.c_int, [C.POINTER(KernelT), _P, _I64, C.c_double, _P, _I64, _I64, _P,
                                   _I64, _P, _I64, _P]),
    "pgp_trsv": (C.c_int, [C.c_int, _P, _I64, _I64, _P, _P, _P, C.c_int, _P]),
    "pgp_trsm": (C.c_int, [C.c_int, _P, _I64, _I64, _P, _P, _I64, _I64, _P]),
    "pgp_workspace_size": (_I64, [_I64, _I32, _I32]),
    "pgp_predict_workspace_size": (_I64, [_I64, _I64, _I32]),
    "pgp_grad_workspace_size": (_I64, [_I64, _I32]),
    "pgp_laplace_fit": (C.c_int, [C.POINTER(ProblemT), C.POINTER(KernelT), C.POINTER(LikelihoodT),
                                  C.POINTER(FitOptionsT), C.POINTER(FitResultT), _P, _I64, _P]),
    "pgp_noalloc_scratch_bytes": (_I64, []),
    "pgp_fit_graph_stats": (C.c_int, [C.POINTER(C.c_int64)] * 3),
    "pgp_laplace_fit_noalloc": (C.c_int, [C.POINTER(ProblemT), C.POINTER(KernelT),
                                          C.POINTER(LikelihoodT), C.POINTER(FitOptionsT),
                                          C.POINTER(FitResultT), _P]),
    "pgp_laplace_set_state": (C.c_int, [C.POINTER(ProblemT), C.POINTER(KernelT),
                                        C.POINTER(LikelihoodT), _I32, _P, _P,
                                        C.POINTER(FitResultT), _P, _I64, _P]),
    "pgp_predict": (C.c_int, [C.POINTER(KernelT), _P, _I64, _P, _P, _P, _I64, _P, _I64, _P]),
    "pgp_predict_covariance": (C.c_int, [C.POINTER(KernelT), _P, _I64, _P, _I64, _P, _I64, _P,
                                         _I64, _P]),
    "pgp_laplace_grad": (C.c_int, [C.POINTER(KernelT), C.POINTER(LikelihoodT), _P, _P, _I64, _P,
                                   _I64, _P]),
    "pgp_vb_map": (C.c_int, [C.POINTER(ProblemT), C.POINTER(KernelT), C.POINTER(LikelihoodT),
                             C.c_int32, _P, _P, _I64, _P]),
    "pgp_vb_grad": (C.c_int, [C.POINTER(KernelT), C.POINTER(LikelihoodT), _P, _P, _I64, _P,
                                   _I64, _P]),
    "pgp_vb_fit": (C.c_int, [C.POINTER(ProblemT), C.POINTER(KernelT), C.POINTER(LikelihoodT),
                             C.POINTER(FitOptionsT), C.POINTER(FitResultT), _P, _I64, _P, _I64,
                             _P]),
    "pgp_batcher_table_size": (_I64, [_I32]),
    "pgp_batcher_create": (C.c_int, [_I32, C.POINTER(FitOptionsT), _P, _I64, _P,
                                     C.POINTER(C.c_void_p)]),
    "pgp_batcher_submit": (C.c_int, [_P, C.POINTER(BatchItemT), _I64]),
    "pgp_batcher_step": (C.c_int, [_P, C.POINTER(C.c_int64), _I32, C.POINTER(C.c_int32)]),
    "pgp_batcher_launch": (C.c_int, [_P]),
    "pgp_batcher_finish": (C.c_int, [_P, C.POINTER(C.c_int64), _I32, C.POINTER(C.c_int32)]),
    "pgp_batcher_in_flight": (_I32, [_P]),
    "pgp_batcher_stats": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "pgp_batcher_launched": (_I32, [_P]),
    "pgp_batcher_ready": (_I32, [_P]),
    "pgp_batcher_destroy": (C.c_int, [_P]),
}

EXPORTED = tuple(_SIGNATURES)

_LIB = None


def load(path: os.PathLike | None = None) -> C.CDLL:
    """Load the library and bind every exported symbol (no GPU needed)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python -m probit_amd.build` "
            "(there is no non-HIP fallback)")
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


_TABLE_LOADED = False


def lib() -> C.CDLL:
    """The library, for a compute call: requires a visible GPU.  The first
    call applies the persisted schedule table for this GPU's architecture
    (probit_amd.autotune.load: PGP_TUNE_FILE, else the package's table), so
    every path of the process factors with the same schedule."""
    global _TABLE_LOADED
    if not torch.cuda.is_available():
        raise RuntimeError("probit_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    L = load()
    if not _TABLE_LOADED:
        _TABLE_LOADED = True
        from . import autotune
        autotune.load()
    return L


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().pgp_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (status {rc}): {msg}")


def info_bits(rc: int, what: str) -> int:
    """Entry points that act on a resident factor return a negative status
    for a bad call (raised here) and positive INFO_* bits when their numbers
    are invalid (returned)."""
    if rc < 0:
        check(rc, what)
    return int(rc)


def stream_handle(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    return t.data_ptr()


def workspace(nbytes: int, device) -> torch.Tensor:
    """Caller-owned device workspace (uint8, 256-byte aligned by torch)."""
    return torch.empty(int(nbytes), dtype=torch.uint8, device=device)
