"""Writes newton_straight21_moment.npz: the oracle's Newton load path on the
21 x 21 straight-crease sheet (rho = 3.75 h, quadratic RKPM, quartic window,
elastic hinge 1e-2 E t^3) under the edge-moment load case of
workloads.with_edge_moment_load: 20 proportional increments, one LineSearch
each (oracle/newton.py, a restatement of ContinuationSolver::LineSearch),
tolerances from workloads.newton_tolerances.

Stored: the converged displacement, the Newton iterations of each increment,
and the first increment's state. Run from the repository root:
    python tests/golden/make_newton_fixture.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "curvami-official_amd")]

import newton as N  # noqa: E402
import oracle as O  # noqa: E402
import workloads as W  # noqa: E402

RHO_FACTOR = 3.75


def main():
    sh = W.with_edge_moment_load(W.config1(rho_factor=RHO_FACTOR))
    p = O.plate_from_sheet(sh, nthreads=os.cpu_count() or 1)
    prob = N.OracleProblem(p, nthreads=os.cpu_count() or 1)
    fext = np.abs(p.external_force()).max()
    tol = W.newton_tolerances(fext)
    t0 = time.time()
    x1 = np.zeros(prob.num_dofs)
    ok1, it1 = N.line_search(prob, 1.0 / W.NEWTON_INCREMENTS, x1, **tol)
    x, conv, its = N.load_path(prob, W.NEWTON_INCREMENTS, 1.0, **tol)
    print(f"{time.time() - t0:.1f} s, converged {sum(conv)}/{len(conv)}, iterations {its}")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "newton_straight21_moment.npz")
    np.savez_compressed(out, x=x, iterations=np.array(its), converged=np.array(conv), x_first=x1,
                        tol_force=tol["tol_force"], tol_energy=tol["tol_energy"], fext_inf=fext)
    print("wrote", out)


if __name__ == "__main__":
    main()
