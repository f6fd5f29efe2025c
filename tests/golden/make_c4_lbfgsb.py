"""Write tests/golden/c4_lbfgsb_probit_n8192.npz: the hyper-parameter loop of
the evidence-gradient configuration (probit J = 3, N = 8192, D = 4, EQ,
X ~ N(0, I), seed 3) driven by the ORACLE -- objective_LA at the FP64 Laplace
fixed point and its closed-form total derivative -- through the same SciPy
L-BFGS-B wrapper the product uses (probit_amd.optimize.minimise_l_bfgs_b, at
most 50 iterations), from the start ((0.5 l, 0.7 s^2), (1.3 sigma, cutpoints)).
The GPU test replays the loop with the HIP objective and adjoint and compares
every evaluation.  Run from the repository root:
    python tests/golden/make_c4_lbfgsb.py          (about 10-30 minutes on 8 cores)
"""

from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from probit_amd import data as DG  # noqa: E402  (host-side data generator)
from probit_amd.optimize import minimise_l_bfgs_b  # noqa: E402  (SciPy wrapper)
from tests.helpers import oracle_value_and_grad  # noqa: E402

N = 8192
OUT = os.path.join(ROOT, "tests", "golden", f"c4_lbfgsb_probit_n{N}.npz")


def main():
    p = DG.evidence_gradient(seed=3, N=N, lik="probit")
    f = oracle_value_and_grad(p.X, p.y, "EQ", "probit")
    noise, cut = p.likelihood_parameters()
    start = ((0.5 * p.lengthscale, 0.7 * p.variance), (1.3 * noise, cut))
    evals = []

    def logged(params):
        t0 = time.perf_counter()
        v, g = f(params)
        (ls, var), (s, c) = params
        evals.append([float(ls), float(var), float(s), *np.asarray(c)[1:-1].tolist(), v])
        print(json.dumps({"eval": len(evals), "value": v, "seconds":
                          round(time.perf_counter() - t0, 1)}), flush=True)
        return v, g

    opt, hist = minimise_l_bfgs_b(logged, start, iters=50)
    (ls, var), (s, c) = opt
    np.savez(OUT, evals=np.array(evals), history=np.array(hist),
             opt=np.array([ls, var, s, *np.asarray(c)[1:-1]]),
             start=np.array([start[0][0], start[0][1], start[1][0], *cut[1:-1]]))
    print(f"wrote {OUT}: {len(hist)} evaluations, objective {hist[0]:.6f} -> {min(hist):.6f}")


if __name__ == "__main__":
    main()
