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
                "--open-loop", "--cpu-seconds", "--copy-calibration"):
        assert opt in r.stdout, opt
