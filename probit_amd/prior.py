"""GP prior samples on the GPU: f = L_K z, L_K = chol(K + jitter I), K = kernel(X, X).

The reference draws its synthetic latents this way in its data generators
(examples/regression.py:52-58, examples/classification.py:227-234):
``L_K = jnp.linalg.cholesky(kernel(X) + jitter * I); f = L_K @ z``.  Here the
Gram matrix, the FP64 factor and the triangular product run as HIP kernels
(`pgp_prior_sample`); z may be given (any generator) or drawn with torch's
generator on the device.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .kernels import Kernel, _device_matrix


def sample_prior(kernel: Kernel, X, z=None, n_samples: int = 1, jitter: float = 1e-10,
                 generator: torch.Generator | None = None) -> torch.Tensor:
    """Samples of the GP prior at X (float64, on the GPU).

    z: standard-normal draws, shape (N,) or (N, S); if None, (N, n_samples) are
    drawn on the device.  Returns f with the shape of z ((N,) for a vector).
    A kernel matrix that is not positive definite after the jitter gives NaN
    samples, as the reference's Cholesky does."""
    lib = _lib.lib()
    Xd = _device_matrix(X)
    N, D = Xd.shape
    if z is None:
        zt = torch.randn((int(n_samples), N), dtype=torch.float64, device=Xd.device,
                         generator=generator)
        vector = False
    else:
        zz = z if isinstance(z, torch.Tensor) else torch.as_tensor(np.asarray(z))
        zz = zz.to(device=Xd.device, dtype=torch.float64)
        vector = zz.dim() == 1
        if vector:
            zz = zz[:, None]
        if zz.dim() != 2 or zz.shape[0] != N:
            raise ValueError(f"z must have {N} rows (got shape {tuple(zz.shape)})")
        zt = zz.t().contiguous()  # (S, N): column-major N x S, one sample per column
    S = zt.shape[0]
    ft = torch.empty_like(zt)
    ws = _lib.workspace(lib.pgp_prior_sample_workspace_size(N), Xd.device)
    rc = lib.pgp_prior_sample(kernel.to_struct(D), Xd.data_ptr(), N, float(jitter),
                              zt.data_ptr(), S, N, ft.data_ptr(), N, ws.data_ptr(), ws.numel(),
                              _lib.stream_handle(Xd.device))
    _lib.check(rc, "pgp_prior_sample")
    f = ft.t()
    return f[:, 0] if vector else f
