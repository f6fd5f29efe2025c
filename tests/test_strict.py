"""strict=True: a fit whose status says its numbers are not valid raises
FloatingPointError; by default the reference's behaviour (NaN, no
exception) is kept and the status stays in last_result.  CPU-only: the
status mapping and the switch, without a fit."""

from __future__ import annotations

import inspect
import math

import pytest


def test_numerical_failure_names_each_failure_bit():
    from probit_amd import _lib
    from probit_amd.approximators import numerical_failure
    assert numerical_failure(0) is None
    assert numerical_failure(_lib.INFO_ESCALATED) is None  # informational only
    assert "pivot" in numerical_failure(_lib.INFO_NOT_PD)
    assert "negative precision" in numerical_failure(_lib.INFO_NEGATIVE_W)
    assert "maxiter" in numerical_failure(_lib.INFO_MAXITER)
    assert "bounded wait" in numerical_failure(_lib.INFO_STALLED)
    assert "non-finite objective" in numerical_failure(0, math.nan)
    both = numerical_failure(_lib.INFO_NOT_PD | _lib.INFO_MAXITER | _lib.INFO_ESCALATED)
    assert "pivot" in both and "maxiter" in both


@pytest.mark.parametrize("cls", ["LaplaceGP", "VBGP"])
def test_strict_switch(cls):
    import probit_amd as P
    from probit_amd import _lib
    from probit_amd.approximators import Approximator
    klass = getattr(P, cls)
    assert issubclass(klass, Approximator)  # LaplaceGP / VBGP forward their arguments
    sig = inspect.signature(Approximator.__init__)
    assert sig.parameters["strict"].default is False
    assert sig.parameters["strict"].kind is inspect.Parameter.KEYWORD_ONLY
    gp = object.__new__(klass)  # no device state needed for the status check
    bad = _lib.FitResultT()
    bad.status = _lib.INFO_NOT_PD
    ok = _lib.FitResultT()
    ok.status = _lib.INFO_ESCALATED
    gp.strict = False
    gp._check_status(bad, "fit")  # the reference's behaviour: no exception
    gp.strict = True
    gp._check_status(ok, "fit")
    with pytest.raises(FloatingPointError, match="non-positive pivot"):
        gp._check_status(bad, "fit")
