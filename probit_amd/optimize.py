"""L-BFGS-B over a GP parameter tree: the hyper-parameter loop of the
reference's examples (varz ``minimise_l_bfgs_b`` at examples/regression.py:157
and examples/classification.py:454), driven by the HIP objective and its
hand-written adjoint.

Parameters are the reference's nested tuples, e.g. ``((lengthscale,
signal_variance), (noise_std, cutpoints))``.  Leaves with only finite
entries are positive and optimised in log space (varz's ``positive``
variables); a leaf with infinite ends is a cutpoint vector: its finite
entries are optimised as the first cutpoint plus log-gaps, which keeps them
ascending (``check_cutpoints``, probit/utilities.py:263-323, rejects any
other order).
"""

from __future__ import annotations

import numpy as np
from scipy.optimize import minimize

__all__ = ["minimise_l_bfgs_b"]


def _leaves(tree, out):
    if isinstance(tree, (tuple, list)):
        return type(tree), [_leaves(t, out) for t in tree]
    out.append(np.asarray(tree, dtype=np.float64))
    return None


def _build(spec, leaves, it):
    if spec is None:
        v = leaves[next(it)]
        return float(v) if v.ndim == 0 else v
    typ, kids = spec
    return typ(_build(k, leaves, it) for k in kids)


class _Packing:
    """Free coordinates x <-> parameter leaves, with the chain rule."""

    def __init__(self, leaves):
        self.leaves = [lv.copy() for lv in leaves]
        self.kinds = []
        for lv in self.leaves:
            fin = np.isfinite(lv)
            if fin.all():
                if np.any(lv <= 0):
                    raise ValueError("positive parameters must be > 0")
                self.kinds.append(("log", None))
            else:
                self.kinds.append(("ordered", np.flatnonzero(fin)))

    def pack(self):
        xs = []
        for lv, (kind, idx) in zip(self.leaves, self.kinds):
            if kind == "log":
                xs.append(np.log(lv).ravel())
            else:
                b = lv[idx]
                xs.append(np.concatenate([b[:1], np.log(np.diff(b))]))
        return np.concatenate(xs) if xs else np.zeros(0)

    def unpack(self, x):
        out, o = [], 0
        for lv, (kind, idx) in zip(self.leaves, self.kinds):
            if kind == "log":
                n = lv.size
                out.append(np.exp(x[o:o + n]).reshape(lv.shape))
            else:
                n = idx.size
                v = lv.copy()
                v[idx] = x[o] + np.concatenate([[0.0], np.cumsum(np.exp(x[o + 1:o + n]))])
                out.append(v)
            o += n
        return out

    def grad_x(self, x, grads):
        gx, o = [], 0
        for lv, (kind, idx), g in zip(self.leaves, self.kinds, grads):
            g = np.asarray(g, dtype=np.float64)
            if kind == "log":
                n = lv.size
                gx.append((g.ravel() * np.exp(x[o:o + n])))
            else:
                n = idx.size
                gb = g[idx]
                # b_j = x0 + sum_{m < j} exp(x_{m+1}):  d/dx0 = sum gb,
                # d/dx_{m+1} = exp(x_{m+1}) * sum_{j > m} gb_j
                tail = np.cumsum(gb[::-1])[::-1]
                gx.append(np.concatenate([[gb.sum()], np.exp(x[o + 1:o + n]) * tail[1:]]))
            o += n
        return np.concatenate(gx) if gx else np.zeros(0)


def minimise_l_bfgs_b(value_and_grad, parameters, iters=50, callback=None):
    """Minimise ``value_and_grad(parameters) -> (value, grads)`` (grads shaped
    like ``parameters``, e.g. ``LaplaceGP.value_and_grad()``) with SciPy's
    L-BFGS-B for at most ``iters`` iterations.  Returns
    ``(parameters_opt, history)`` with the objective value of every
    evaluation in ``history``."""
    flat = []
    spec = _leaves(parameters, flat)
    pk = _Packing(flat)
    history = []

    def fun(x):
        params = _build(spec, pk.unpack(x), iter(range(len(flat))))
        val, grads = value_and_grad(params)
        gl = []
        _leaves(grads, gl)
        history.append(float(val))
        if not np.isfinite(val):
            return np.inf, np.zeros_like(x)
        return float(val), pk.grad_x(x, gl)

    res = minimize(fun, pk.pack(), jac=True, method="L-BFGS-B",
                   options={"maxiter": int(iters)}, callback=callback)
    return _build(spec, pk.unpack(res.x), iter(range(len(flat)))), history
