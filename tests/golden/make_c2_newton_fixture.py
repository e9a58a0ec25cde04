"""Writes newton_arc201_moment.npz: the oracle's Newton load path on the
201 x 201 arc-crease sheet (arc r = 0.4 about (0.5, -0.1) as 256 segments,
rho = 3.75 h, quadratic RKPM, quartic window, elastic hinge 1e-2 E t^3) under the
edge-moment load case of workloads.with_edge_moment_load: 20 proportional
increments, one LineSearch each (oracle/newton.py, the restatement of
ContinuationSolver::LineSearch on the oracle's residual, tangent, constraint
rows and load), tolerances from workloads.newton_tolerances. The KKT systems
(121,203 DOFs, 774 constraint rows) are solved by the library's host KKT solver
(libcurvami_kkt; bit-reproducible), which is not under test here: what the
fixture pins is the load path of the oracle's discrete problem.

Stored after every increment (so that a partial run is usable): the Newton
iterations and convergence of each increment done, the states after increments
1 and 10 and after the last one done. Run from the repository root (hours of
CPU time):
    python tests/golden/make_c2_newton_fixture.py [threads]
"""
import os
import sys
import time

import numpy as np
import scipy.sparse as sp

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "curvami-official_amd")]

import curvami_gpu as G  # noqa: E402
import newton as N  # noqa: E402
import oracle as O  # noqa: E402
import workloads as W  # noqa: E402

RHO_FACTOR = 3.75
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "newton_arc201_moment.npz")


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    sh = W.with_edge_moment_load(W.config2(rho_factor=RHO_FACTOR))
    t0 = time.time()
    p = O.plate_from_sheet(sh, nthreads=threads)
    prob = N.OracleProblem(p, nthreads=threads)
    fext = np.abs(p.external_force()).max()
    tol = W.newton_tolerances(fext)
    s = G.KKT(threads)

    def kkt(K, Jd, rhs):
        K = sp.csc_matrix(K)
        J = sp.csr_matrix(Jd)
        rows = np.repeat(np.arange(J.shape[0], dtype=np.int32), np.diff(J.indptr))
        s.set_pattern(K.indptr, K.indices, J.shape[0], rows, J.indices)
        s.factorize(K.data, J.data)
        return s.solve(rhs)

    print(f"set-up {time.time() - t0:.1f} s: {sh.nn} nodes, {sh.face.n} face GPs", flush=True)
    its, conv, keep = [], [], {}

    def on_increment(k, x, it):
        its.append(it); conv.append(it > 0)
        if k in (1, 10):
            keep[k] = x.copy()
        np.savez_compressed(OUT, iterations=np.array(its), converged=np.array(conv), x_last=x,
                            last=k, x_first=keep.get(1, x), x_10=keep.get(10, np.zeros(0)),
                            tol_force=tol["tol_force"], tol_energy=tol["tol_energy"], fext_inf=fext)
        print(f"increment {k}: {it} iterations, {time.time() - t0:.0f} s, {s.stats()}", flush=True)

    N.load_path(prob, W.NEWTON_INCREMENTS, 1.0, on_increment=on_increment, kkt=kkt, **tol)
    print(f"{time.time() - t0:.1f} s, converged {sum(conv)}/{len(conv)}, iterations {its}; wrote {OUT}")


if __name__ == "__main__":
    main()
