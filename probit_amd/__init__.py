"""probit_amd: MI355X (gfx950) HIP implementation of probit's GP posterior path.

Drop-in for ``probit.approximators.LaplaceGP`` / ``VBGP``, the likelihood
utilities of ``probit.utilities`` and the mlkernels kernel expressions the
reference's priors use.  All compute runs in ``libprobit_amd.so`` (hand-written
HIP kernels behind a C ABI, ``include/probit_amd.h``); build it with
``python -m probit_amd.build``.
"""

from .approximators import Approximator, LaplaceGP, VBGP
from .kernels import EQ, Kernel, Matern12
from .optimize import minimise_l_bfgs_b
from .utilities import (BOUNDS, CutpointValueError, InvalidKernel, calculate_metrics,
                        check_cutpoints, check_data,
                        grad_log_probit_likelihood, h, hessian_log_probit_likelihood,
                        log_gaussian_likelihood, log_probit_likelihood, ndtr, norm_cdf,
                        norm_logpdf, norm_pdf, norm_z_logpdf, norm_z_pdf, probit,
                        probit_likelihood, probit_predictive_distributions)

__all__ = [
    "Approximator", "LaplaceGP", "VBGP", "EQ", "Matern12", "Kernel", "BOUNDS",
    "CutpointValueError", "InvalidKernel", "check_cutpoints", "check_data",
    "grad_log_probit_likelihood", "hessian_log_probit_likelihood", "h",
    "log_gaussian_likelihood", "log_probit_likelihood", "probit_predictive_distributions",
    "ndtr", "norm_cdf", "norm_pdf", "norm_z_pdf", "norm_logpdf", "norm_z_logpdf", "probit",
    "probit_likelihood", "minimise_l_bfgs_b", "calculate_metrics",
]
