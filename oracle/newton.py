"""Oracle Newton load path -- TEST INFRASTRUCTURE ONLY (imported by tests/).

A direct restatement of ContinuationSolver::LineSearch
(core/include/curvami/solver/ContinuationSolver.h:185-322) with its
InitializeTopology / filterRows (:105-164) and ConstraintProjector
(core/include/curvami/solver/ConstraintProjector.h:19-55), written against the
oracle plate (oracle.py) so that the whole load path -- residual, tangent,
constraint rows, external load -- is the CPU restatement. The KKT system of
MumpsSymmetricSolver (core/include/curvami/solver/MumpsSymmetricSolver.h:46-55)
is solved with scipy's sparse LU (spsolve); the projector's Gram system with a
dense solve.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla


class OracleProblem:
    """PlateProblem (plate/PlateProblem.h:52-91) on the oracle plate."""

    def __init__(self, plate, flag_rig=False, nthreads=1, exact_tangent=False):
        self.p = plate
        self.flag_rig = flag_rig
        self.exact_tangent = exact_tangent
        self.nthreads = nthreads
        self.num_dofs = 3 * plate.nn
        self._fext = None

    def assemble(self, x):
        outer, inner, val, F = self.p.assemble(x, flag_rig=self.flag_rig, exact_tangent=self.exact_tangent,
                                               nthreads=self.nthreads)
        n = self.num_dofs
        return sp.csc_matrix((val, inner, outer), shape=(n, n)), F

    def internal_force(self, x):
        return self.p.internal_force(x, flag_rig=self.flag_rig, with_tangent=False, nthreads=self.nthreads)

    def external_force(self, x):
        if self._fext is None:
            self._fext = self.p.external_force()
        return self._fext

    def load_stiffness(self, x):
        return None

    def constraints(self, x):
        c, r, col, v = self.p.constraints(x, flag_rig=self.flag_rig)
        return c, sp.coo_matrix((v, (r, col)), shape=(c.size, self.num_dofs)).tocsr()


def _sym(K):
    # selfadjointView<Upper>: the stored lower halves of the diagonal 3 x 3 node
    # blocks are ignored
    return sp.triu(K) + sp.triu(K, 1).T


def _independent_rows(J):
    if J.shape[0] == 0:
        return np.zeros(0, np.int64)
    R, perm = sla.qr(J.toarray().T, mode="r", pivoting=True)
    d = np.abs(np.diag(R))
    tol = max(J.shape) * np.finfo(float).eps * d.max()
    return np.sort(perm[:d.size][d > tol])


def _projector(Jd):
    """Returns (corrector(c), project(R)) for a dense J_ind."""
    if Jd.shape[0] == 0:
        return (lambda c, ns: np.zeros(ns)), (lambda R: R)
    G = Jd @ Jd.T
    return (lambda c, ns: Jd.T @ np.linalg.solve(G, -c)), (lambda R: R - Jd.T @ np.linalg.solve(G, Jd @ R))


def line_search(prob, lam, d, tol_force=1e-8, tol_energy=1e-8, tol_geom=1e-12, div=1e7, max_iters=500):
    """Returns (converged, iterations); d is updated in place (rolled back on failure)."""
    ns = d.size
    d0 = d.copy()
    c, J = prob.constraints(d)
    ind = _independent_rows(J)
    energy = 1.0
    for j in range(1, max_iters + 1):
        c, J = prob.constraints(d)
        ci, Jd = c[ind], J[ind].toarray()
        R = prob.internal_force(d) - lam * prob.external_force(d)
        corr, proj = _projector(Jd)
        mf = np.abs(proj(R)).max()
        ge = np.abs(ci).max() if ci.size else 0.0
        if j > 1 and mf <= tol_force and energy <= tol_energy and ge <= tol_geom:
            return True, j
        if mf >= div or np.isnan(mf):
            d[:] = d0
            return False, -1
        K, _ = prob.assemble(d)
        xb = corr(ci, ns)
        Ks = _sym(K)
        A = sp.bmat([[Ks, sp.csr_matrix(Jd).T], [sp.csr_matrix(Jd), None]], format="csc") if ci.size else Ks.tocsc()
        rhs = np.concatenate([-(R + Ks @ xb), np.zeros(ci.size)])
        dx = spla.spsolve(A, rhs)[:ns]
        alpha = 1.0
        for _ in range(6):
            dt = d + xb + alpha * dx
            Rt = prob.internal_force(dt) - lam * prob.external_force(dt)
            ct, Jt = prob.constraints(dt)
            _, pt = _projector(Jt[ind].toarray())
            if np.abs(pt(Rt)).max() <= 1.05 * mf:
                break
            alpha *= 0.5
        energy = abs(np.dot(xb + alpha * dx, R))
        d[:] = dt
    d[:] = d0
    return False, -1


def load_path(prob, increments=20, lam_total=1.0, **tol):
    x = np.zeros(prob.num_dofs)
    conv, its = [], []
    for k in range(1, increments + 1):
        ok, it = line_search(prob, lam_total * k / increments, x, **tol)
        conv.append(ok); its.append(it)
        if not ok:
            break
    return x, conv, its
