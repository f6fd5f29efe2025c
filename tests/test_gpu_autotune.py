"""The schedule table of the blocked Cholesky and its autotuner
(probit_amd.autotune; the factorisation it tunes is the reference's
B.cholesky, probit/implicit/Laplace.py:24): a tuned super-step width is the
one the factorisation uses (its rounding shows in FP32), COLUMN_MAX /
FUSE_MAX / COLUMN_MERGE change the launches but not one bit of the results
(the merged super-step kernel against one launch per column, in every
precision, single fits and batcher mixes), results stay
the same alone and in a mix under per-class entries, and a tuning run
persists a table that loads back."""

from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import make_prior, small_problem

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_table():
    from probit_amd import autotune
    saved = autotune.table()
    yield autotune
    for p, entries in saved.items():
        for c, e in enumerate(entries):
            autotune.set_entry(p, c, e)


def _fit(N, mode="fp32", seed=90):
    import probit_amd as P
    X, y, theta, lp, _ = small_problem(seed, N, 3, "probit", "Matern12", noise=0.5)
    gp = P.LaplaceGP(data=(X, y), prior=make_prior("Matern12"),
                     log_likelihood=P.log_probit_likelihood, precision_mode=mode)
    w, _ = gp.approximate_posterior((theta, lp))
    return w.cpu().numpy(), gp.last_result.iterations


def _panel_launches():
    import ctypes as C

    from probit_amd import _lib
    lib = _lib.lib()
    n = C.c_int64()
    d = C.c_double()
    _lib.check(lib.pgp_timer_read(_lib.TIMER_PANEL, C.byref(n), C.byref(d), C.byref(d),
                                  C.byref(d)), "timer")
    return n.value


def test_table_entry_is_used_and_launch_choices_are_result_neutral(restore_table):
    from probit_amd import _lib
    at = restore_table
    lib = _lib.lib()
    base = {"superstep": 8, "column_max": 512, "fuse_max": 512, "column_merge": 1}
    at.set_entry("fp32", 0, base)
    lib.pgp_timer_reset()
    lib.pgp_timer_enable(1 << _lib.TIMER_PANEL)
    w_a, it_a = _fit(960)
    n_a = _panel_launches()
    lib.pgp_timer_reset()
    at.set_entry("fp32", 0, dict(base, column_max=0, fuse_max=0))
    w_c, it_c = _fit(960)
    n_c = _panel_launches()
    lib.pgp_timer_enable(0)
    # another launch schedule, the same bits
    assert n_a != n_c
    assert np.array_equal(w_a, w_c) and it_a == it_c
    # another super-step width: other FP32 rounding, so the table is used
    at.set_entry("fp32", 0, dict(base, superstep=3))
    w_b, _ = _fit(960)
    assert not np.array_equal(w_a, w_b)
    assert np.max(np.abs(w_a - w_b)) <= 1e-6 * np.max(np.abs(w_a))
    assert at.get("fp32", 0) == dict(base, superstep=3)


def test_class_entries_keep_results_the_same_alone_and_in_a_mix(restore_table):
    """Per-class super-step widths (class 0: 4, class 1: 6): each problem's
    factor follows its own class, so a mix through the batcher gives every
    problem the bits of its single fit."""
    import torch

    import probit_amd as P
    from probit_amd.serving import ServingEngine
    at = restore_table
    at.set_entry("fp32", 0, {"superstep": 4, "column_max": 512, "fuse_max": 512})
    at.set_entry("fp32", 1, {"superstep": 6, "column_max": 256, "fuse_max": 0, "column_merge": 0})
    sizes = (300, 1900, 900, 1300)
    probs = [small_problem(100 + k, N, 2, "probit", "EQ", noise=0.5) for k, N in enumerate(sizes)]
    eng = ServingEngine(precision="fp32", capacity=8)
    for tag, (X, y, theta, lp, _) in enumerate(probs):
        assert eng.submit(tag, X, y, "EQ", theta[0], theta[1], "probit", lp[0], lp[1])
    out = {}
    while eng.in_flight:
        for c in eng.step():
            out[c.tag] = c
    eng.close()
    for tag, (X, y, theta, lp, _) in enumerate(probs):
        gp = P.LaplaceGP(data=(X, y), prior=make_prior("EQ"),
                         log_likelihood=P.log_probit_likelihood, precision_mode="fp32")
        w, _ = gp.approximate_posterior((theta, lp))
        assert torch.equal(w, out[tag].weight), (tag, sizes[tag])
        assert gp.last_result.iterations == out[tag].iterations


def test_tuning_run_persists_a_table_that_loads_back(restore_table, tmp_path):
    at = restore_table
    path = tmp_path / "tune.json"
    res = at.tune(precisions=("fp32",), classes=(0,),
                  candidates={"superstep": (4, 8), "column_max": (0, 512), "fuse_max": (512,),
                              "column_merge": (2,)},
                  window=24, capacity=16, reps=1, path=path)
    best = res["fp32"][0]["entry"]
    assert best["superstep"] in (4, 8) and best["column_max"] in (0, 512)
    assert best["column_merge"] == 2
    assert len(res["fp32"][0]["trials"]) >= 3
    assert at.get("fp32", 0) == best
    at.set_entry("fp32", 0, {"superstep": 2, "column_max": 7, "fuse_max": 9, "column_merge": 0})
    assert at.load(path)
    assert at.get("fp32", 0) == best


@pytest.mark.parametrize("mode", ["fp64", "fp32", "bf16"])
def test_merged_superstep_kernel_equals_per_column_launches(restore_table, mode):
    """The two merged forms of a column-form super-step (its diagonal block in
    one k_superstep launch and the rows below in one k_panel_rows launch; all
    of it in one k_superstep launch) against one k_column launch per block
    column: the same factor bits, single fits (Gaussian and probit) and a
    batcher mix of sizes in both classes, with fewer launches."""
    import torch

    from probit_amd import _lib
    from probit_amd.serving import ServingEngine
    at = restore_table
    lib = _lib.lib()
    entry = {"superstep": 6, "column_max": 4096, "fuse_max": 512}
    sizes = (1500, 700, 1100, 300, 64, 900)
    probs = [small_problem(140 + k, N, 2, "probit" if k % 2 else "gaussian", "EQ", noise=0.6)
             for k, N in enumerate(sizes)]

    def run(merge):
        for c in range(_lib.TUNE_CLASSES):
            at.set_entry(mode, c, dict(entry, column_merge=merge))
        lib.pgp_timer_reset()
        lib.pgp_timer_enable(1 << _lib.TIMER_PANEL)
        single = [_fit(N, mode=mode, seed=150 + N) for N in (700, 1300)]
        eng = ServingEngine(precision=mode, capacity=8)
        for tag, (X, y, theta, lp, _) in enumerate(probs):
            lik = "probit" if tag % 2 else "gaussian"
            assert eng.submit(tag, X, y, "EQ", theta[0], theta[1], lik, lp[0],
                              lp[1] if lik == "probit" else None)
        out = {}
        while eng.in_flight:
            for c in eng.step():
                out[c.tag] = c
        eng.close()
        launches = _panel_launches()
        lib.pgp_timer_enable(0)
        return single, out, launches

    print(f"merged superstep {mode}", flush=True)
    s0, b0, n0 = run(0)
    for merge in (1, 2):
        s1, b1, n1 = run(merge)
        assert n1 < n0, (merge, n1, n0)
        for (w1, it1), (w0, it0) in zip(s1, s0):
            assert np.array_equal(w1, w0) and it1 == it0, merge
        assert sorted(b1) == sorted(b0) == list(range(len(sizes)))
        for t in b1:
            for name in ("mean", "var", "weight", "precision"):
                a, b = getattr(b1[t], name), getattr(b0[t], name)
                if a is None:
                    assert b is None
                    continue
                assert torch.equal(a, b), (mode, merge, t, sizes[t], name)
            assert b1[t].iterations == b0[t].iterations and b1[t].status == b0[t].status
            assert b1[t].status & ~_lib.INFO_ESCALATED == 0, (t, b1[t].status)
