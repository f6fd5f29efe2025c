"""Host logic of the hyper-parameter loop (probit_amd.optimize): the packing
of a parameter tree into free coordinates, its chain rule, and convergence
on a function with a known minimum (no GPU)."""

import numpy as np

from probit_amd.optimize import _Packing, minimise_l_bfgs_b


def test_packing_round_trip_and_chain_rule():
    leaves = [np.array(0.8), np.array([0.5, 2.0]), np.array(0.6),
              np.array([-np.inf, -0.3, 0.4, 1.5, np.inf])]
    pk = _Packing(leaves)
    x = pk.pack()
    for a, b in zip(pk.unpack(x), leaves):
        assert np.array_equal(a, b) or np.allclose(a[np.isfinite(b)], b[np.isfinite(b)])
    rng = np.random.default_rng(0)
    coef = [np.where(np.isfinite(lv), rng.normal(size=lv.shape), 0.0) for lv in leaves]

    def F(xv):
        return sum(float(np.sum(np.where(np.isfinite(u), c * u, 0.0)))
                   for c, u in zip(coef, pk.unpack(xv)))

    g = pk.grad_x(x, coef)
    fd = np.array([(F(x + 1e-6 * e) - F(x - 1e-6 * e)) / 2e-6 for e in np.eye(x.size)])
    assert np.max(np.abs(g - fd)) < 1e-8


def test_minimise_finds_known_minimum_and_keeps_cutpoints_ordered():
    target_ls, target_var, target_cut = 1.7, 0.4, np.array([-0.2, 0.9])

    def value_and_grad(params):
        (ls, var), (noise, cut) = params
        c = cut[1:-1]
        val = (np.log(ls / target_ls) ** 2 + np.log(var / target_var) ** 2
               + (noise - 0.3) ** 2 + np.sum((c - target_cut) ** 2))
        gcut = np.zeros_like(cut)
        gcut[1:-1] = 2 * (c - target_cut)
        grads = ((2 * np.log(ls / target_ls) / ls, 2 * np.log(var / target_var) / var),
                 (2 * (noise - 0.3), gcut))
        return val, grads

    start = ((0.5, 2.0), (0.9, np.array([-np.inf, 0.5, 0.6, np.inf])))
    opt, hist = minimise_l_bfgs_b(value_and_grad, start, iters=50)
    (ls, var), (noise, cut) = opt
    assert abs(ls - target_ls) < 1e-4 and abs(var - target_var) < 1e-4
    assert abs(noise - 0.3) < 1e-4
    assert np.allclose(cut[1:-1], target_cut, atol=1e-4)
    assert np.isinf(cut[0]) and np.isinf(cut[-1]) and np.all(np.diff(cut) > 0)
    assert hist[-1] <= hist[0]
