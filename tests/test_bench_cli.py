"""bench.py's command line on the CPU: more than one GPU is refused with a
message (the measured path runs on one GPU; the sharding is described, not
run), and the help lists the options the measurement commands use."""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _run(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True,
                          text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_help_lists_the_measurement_options():
    r = _run("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--gpus", "--precision", "--min-n", "--plain-sample",
                "--open-loop", "--cpu-strata", "--cpu-study", "--profile-window",
                "--copy-calibration", "--kernel-window", "--factor-only"):
        assert opt in r.stdout, opt


def test_cpu_study_runs_without_a_gpu():
    """The CPU-baseline plan on a tiny stream: both oracle variants, all cores
    and one, one JSON line."""
    import json
    r = _run("--cpu-study", "1", "--cpu-strata", "2", "--stream", "6", "--steps", "1",
             "--warmup", "0", "--requests-per-step", "6")
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)["cpu_study"]
    assert set(out["runs"]) == {"lu_all_cores", "lu_1_cores", "bform_all_cores", "bform_1_cores"}
    assert out["runs"]["lu_1_cores"]["cores"] == 1
    assert all(v["posteriors_per_s_median"] > 0 for v in out["runs"].values())
