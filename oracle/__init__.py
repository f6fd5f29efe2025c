"""FP64 CPU oracle for the GP posterior path of bb515/probit.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
reference that results are compared against (or as the timed CPU baseline),
never as the thing measured or shipped.
The product package ``probit_amd`` never imports it.

``gp_oracle`` restates, in NumPy/SciPy float64, the reference's
``probit/approximators.py``, ``probit/implicit/{Laplace,VB,solvers}.py`` and
``probit/utilities.py`` function by function (each function cites the lines it
follows).  ``implicit_torch`` restates the reference's implicit-function
gradient (``fixed_point_layer`` custom VJP) with torch float64 autograd on the
CPU.

Pinning: every oracle function is checked in ``tests/test_oracle.py`` against
something other than its own formula -- the reference's only known-answer test
(``probit/test/test_implicit.py:11-22``), scikit-learn's independent GP
regression (mean, variance, log marginal likelihood and its gradient), SciPy's
normal CDF/PDF, torch autograd derivatives, brute-force loops on tiny
inputs, finite differences and closed-form special cases.  The exceptions are
listed in ``UNPINNED`` below.
"""

# Functions whose value no independent check fixes beyond self-consistency.
UNPINNED = {
    # jaxopt's exact stopping iterate is not reproducible without jaxopt; the
    # oracle compares converged fixed points instead (tolerance 1e-12).
    "fpi_stopping_iterate": "jaxopt FixedPointIteration stopping iterate (jaxopt absent)",
}
