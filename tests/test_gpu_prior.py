"""GP prior samples on the GPU (pgp_prior_sample) against the oracle's
f = chol(K + jitter I) z, the reference data generators' sampling
(examples/regression.py:52-58, examples/classification.py:227-234)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import gp_oracle as O

pytestmark = pytest.mark.gpu


def _kernel(kind, var, ls, period):
    import probit_amd as P
    k = var * (P.EQ() if kind == "EQ" else P.Matern12()).stretch(ls)
    return k.periodic(period) if period is not None else k


@pytest.mark.parametrize("N,D,kind,period,S,jitter", [
    (20, 1, "EQ", 0.5, 0, 1e-10),        # the regression example's prior (vector z)
    (300, 3, "EQ", None, 5, 1e-8),       # ARD lengthscales, padded rows
    (1100, 2, "Matern12", None, 70, 1e-10),  # two sample tiles, 17 block rows
    (4096, 4, "Matern12", None, 3, 1e-6),
])
def test_prior_sample_matches_oracle(N, D, kind, period, S, jitter):
    import torch

    from probit_amd.prior import sample_prior
    rng = np.random.default_rng(N + S)
    X = rng.uniform(-1.0, 1.0, size=(N, D))
    ls = np.exp(rng.uniform(np.log(0.3), np.log(1.5), size=D))
    var = 1.3
    z = rng.standard_normal(N) if S == 0 else rng.standard_normal((N, S))
    print(f"prior sample N={N} D={D} {kind} S={S}", flush=True)
    f = sample_prior(_kernel(kind, var, ls, period), X, z, jitter=jitter)
    torch.cuda.synchronize()
    assert tuple(f.shape) == z.shape
    spec = O.KernelSpec.make(O.EQ if kind == "EQ" else O.MATERN12, var, ls, D,
                             period=None if period is None else [period] * D)
    ref = O.prior_sample(spec, X, z, jitter)
    err = np.max(np.abs(f.cpu().numpy() - ref)) / np.max(np.abs(ref))
    assert err < 1e-8, err


def test_prior_sample_draws_z_on_the_device_and_has_the_prior_covariance():
    """Without z: (N, S) standard-normal draws on the device; the samples'
    covariance is K + jitter I."""
    import torch

    from probit_amd.prior import sample_prior
    rng = np.random.default_rng(11)
    X = rng.uniform(-1.0, 1.0, size=(40, 2))
    kern = _kernel("EQ", 0.9, np.array([0.5, 0.8]), None)
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    f = sample_prior(kern, X, n_samples=100000, jitter=0.05, generator=g)
    assert tuple(f.shape) == (40, 100000)
    fn = f.cpu().numpy()
    C = fn @ fn.T / fn.shape[1]
    K = O.gram(O.KernelSpec.make(O.EQ, 0.9, [0.5, 0.8], 2), X) + 0.05 * np.eye(40)
    assert np.max(np.abs(C - K)) < 0.05 * np.max(np.abs(K))


def test_prior_sample_of_indefinite_kernel_is_nan():
    """K + jitter I not positive definite: NaN samples (as jnp.linalg.cholesky
    gives), no error."""
    import torch

    from probit_amd.prior import sample_prior
    X = np.random.default_rng(12).uniform(-1.0, 1.0, size=(70, 1))
    f = sample_prior(_kernel("EQ", -1.0, 0.5, None), X, np.ones(70), jitter=0.0)
    torch.cuda.synchronize()
    assert torch.isnan(f).all()
