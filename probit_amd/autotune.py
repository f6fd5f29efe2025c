"""The project's own autotuner of the blocked Cholesky schedule.

The factorisation inside every Newton step (the reference's ``B.cholesky``,
probit/implicit/Laplace.py:24, and the LU solve of the Newton system,
probit/implicit/solvers.py:24) runs in super-steps of G 64-column block
columns; COLUMN_MAX decides which super-steps run in the column kernel,
COLUMN_MERGE whether such a super-step is one launch or one launch per block
column, and FUSE_MAX which update launches factor the next diagonal block
inside the launch.  The library keeps one entry (G, COLUMN_MAX, FUSE_MAX,
COLUMN_MERGE) per precision
and size class (padded N up to 1024, or above) in a mutex-guarded table
(``pgp_tune_get`` / ``pgp_tune_set``).  ``tune()`` fills it from timed trials
on the workload it serves: a window of the serving stream (N log-uniform on
[64, 4096], D in 1..8, Gaussian or ordinal probit, EQ or Matern12, 256 test
points) through the asynchronous engine, under each candidate, by coordinate
search per size class (G, then COLUMN_MAX, FUSE_MAX and COLUMN_MERGE), keeps the
fastest and persists the table as JSON per GPU architecture.  ``load()`` applies a persisted table; the package
ships the one tuned on an MI355X (``tune_table.json``), which
``ServingEngine`` loads by default.
"""

from __future__ import annotations

import json
import math
import os
import time
from pathlib import Path

import torch

from . import _lib

__all__ = ["CANDIDATES", "PACKAGE_TABLE", "device_arch", "get", "set_entry", "table", "save",
           "load", "tune"]

CANDIDATES = {"superstep": (2, 3, 4, 6, 8, 12, 16),
              "column_max": (0, 256, 512, 1024, 2048, 4096, 16384),
              "fuse_max": (0, 512), "column_merge": (0, 1, 2)}
PACKAGE_TABLE = Path(__file__).resolve().parent / "tune_table.json"
_FIELDS = ("superstep", "column_max", "fuse_max", "column_merge")
# a persisted entry without a field takes this value
_DEFAULTS = {"column_merge": 2}
_PRECISIONS = ("fp64", "fp32", "bf16")


def device_arch(device=None) -> str:
    """The GPU architecture name a table is keyed by (e.g. gfx950)."""
    props = torch.cuda.get_device_properties(device if device is not None else 0)
    return str(getattr(props, "gcnArchName", props.name)).split(":")[0]


def get(precision: str, size_class: int) -> dict:
    lib = _lib.load()
    e = _lib.TuneEntryT()
    _lib.check(lib.pgp_tune_get(_lib.PRECISIONS[precision], size_class, e), "pgp_tune_get")
    return {f: int(getattr(e, f)) for f in _FIELDS}


def set_entry(precision: str, size_class: int, entry: dict) -> None:
    lib = _lib.load()
    e = _lib.TuneEntryT()
    for f in _FIELDS:
        setattr(e, f, int(entry[f] if f in entry else _DEFAULTS[f]))
    _lib.check(lib.pgp_tune_set(_lib.PRECISIONS[precision], size_class, e), "pgp_tune_set")


def table() -> dict:
    """{precision: [entry of class 0, entry of class 1]} as the library holds it."""
    return {p: [get(p, c) for c in range(_lib.TUNE_CLASSES)] for p in _PRECISIONS}


def save(path=None, arch=None, tab=None) -> Path:
    """Write `tab` (default: the current table) under the architecture key,
    keeping other architectures' entries in the file."""
    path = Path(path) if path is not None else PACKAGE_TABLE
    arch = arch or device_arch()
    data = {}
    if path.exists():
        with open(path) as f:
            data = json.load(f)
    data[arch] = tab if tab is not None else table()
    tmp = path.with_suffix(".tmp")
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
    os.replace(tmp, path)
    return path


def load(path=None, arch=None) -> bool:
    """Apply the persisted entries for this GPU architecture; False if the
    file or the architecture is absent (the library's defaults stay)."""
    path = Path(path) if path is not None else Path(os.environ.get("PGP_TUNE_FILE",
                                                                   PACKAGE_TABLE))
    if not path.exists():
        return False
    with open(path) as f:
        data = json.load(f)
    arch = arch or device_arch()
    tab = data.get(arch)
    if not tab:
        return False
    for p, entries in tab.items():
        for c, e in enumerate(entries):
            set_entry(p, c, e)
    return True


def _stream_window(n: int, seed: int):
    """Device copies of the first n requests of the serving stream."""
    from .data import serving_stream
    out = []
    for r in serving_stream(seed=seed, n=n):
        p = r.problem
        out.append((r.rid, torch.as_tensor(p.X, device="cuda"), torch.as_tensor(p.y, device="cuda"),
                    p.kind, p.lengthscale, p.variance, p.lik, p.noise_std, p.cutpoints,
                    torch.as_tensor(p.Xtest, device="cuda"), p.period))
    return out


def _serve_seconds(window, precision: str, capacity: int, reps: int) -> float:
    """Best wall time of serving the window through the asynchronous engine."""
    from .serving import ServingEngine, serve_async
    best = math.inf
    for _ in range(reps):
        eng = ServingEngine(precision=precision, capacity=capacity)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = serve_async(eng, window)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
        eng.close()
        if len(done) != len(window):
            raise RuntimeError("autotune: the trial window did not complete")
    return best


def tune(precisions=("fp32",), classes=(0, 1), candidates=None, window=1536, capacity=1024,
         reps=2, seed=5, path=None, persist=True, log=None) -> dict:
    """Coordinate search of (superstep, column_max, fuse_max, column_merge) per precision and
    size class.  The objective is the serving workload itself: a window of the
    serving stream (seed `seed`, `window` requests, mixed sizes in both
    classes) served through the asynchronous engine with `capacity` slots;
    each class's entry is searched field by field with the other class's
    fixed.  Sets the best entries in the library, persists them (``path``,
    default the package table) and returns {precision: [per-class result]}."""
    cand = dict(CANDIDATES, **(candidates or {}))
    win = _stream_window(window, seed)
    out = {}
    for prec in precisions:
        out[prec] = []
        _serve_seconds(win, prec, capacity, 1)  # warm-up: kernels loaded, allocator primed
        t_best = _serve_seconds(win, prec, capacity, reps)
        for cls in classes:
            best = get(prec, cls)
            trials = [(dict(best), t_best)]
            for field in _FIELDS:
                for v in cand[field]:
                    if v == best[field]:
                        continue
                    e = dict(best, **{field: v})
                    set_entry(prec, cls, e)
                    t = _serve_seconds(win, prec, capacity, reps)
                    trials.append((e, t))
                    if log:
                        log(f"autotune {prec} class {cls}: {e} {1e3 * t:.2f} ms")
                    if t < t_best:
                        best, t_best = e, t
                set_entry(prec, cls, best)
            out[prec].append({"entry": best, "seconds": t_best,
                              "trials": [[e, round(t, 6)] for e, t in trials]})
    if persist:
        save(path, tab=table())
    return out
