"""The reference's elementwise normal / probit helpers
(probit/utilities.py:18-70,195-229) against SciPy and the oracle, on the CPU."""

import math

import numpy as np
from scipy import stats

from oracle import gp_oracle as O
from tests.helpers import small_problem


def test_normal_helpers_match_scipy():
    from probit_amd import utilities as U
    x = np.linspace(-7.0, 7.0, 301)
    assert np.allclose(U.ndtr(x).numpy(), stats.norm.cdf(x), rtol=1e-12, atol=1e-15)
    assert np.allclose(U.norm_z_pdf(x).numpy(), stats.norm.pdf(x), rtol=1e-13, atol=0)
    assert np.allclose(U.norm_pdf(x, 0.3, 1.7).numpy(), stats.norm.pdf(x, 0.3, 1.7), rtol=1e-13)
    assert np.allclose(U.norm_logpdf(x, 0.3, 1.7).numpy(), stats.norm.logpdf(x, 0.3, 1.7),
                       rtol=1e-13)
    assert np.allclose(U.norm_z_logpdf(x).numpy(), stats.norm.logpdf(x), rtol=1e-13)
    c = U.norm_cdf(np.array([-math.inf, -1.0, 0.0, math.inf])).numpy()
    assert c[0] == 0.0 and c[3] == 1.0 and abs(c[2] - 0.5) < 1e-16


def test_probit_likelihood_matches_oracle():
    from probit_amd import utilities as U
    _, y, _, lp, _ = small_problem(5, 200, 1, "probit", J=4)
    f = np.random.default_rng(6).normal(scale=2.0, size=200)
    Z = U.probit_likelihood(f, y, lp).numpy()
    assert np.allclose(Z, O.probit_Z(f, y, lp), rtol=1e-10, atol=1e-15)
    # the first and last classes use the infinite cutpoints
    assert set(np.unique(y).astype(int)) == {0, 1, 2, 3}
    assert np.allclose(np.log(Z + 1e-10), O.log_probit_likelihood(f, y, lp), rtol=1e-10)
