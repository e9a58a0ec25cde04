"""Oracle Newton load path -- TEST INFRASTRUCTURE ONLY (imported by tests/).

A direct restatement of ContinuationSolver::LineSearch
(core/include/curvami/solver/ContinuationSolver.h:185-322) with its
InitializeTopology / filterRows (:105-164) and ConstraintProjector
(core/include/curvami/solver/ConstraintProjector.h:19-55), written against the
oracle plate (oracle.py) so that the whole load path -- residual, tangent,
constraint rows, external load -- is the CPU restatement. The KKT system of
MumpsSymmetricSolver (core/include/curvami/solver/MumpsSymmetricSolver.h:46-55)
is solved with scipy's sparse LU (spsolve); the projector's Gram system with a
dense solve. arc_length restates ContinuationSolver::ArcLength (:329-538) with
the default step control of core/include/curvami/solver/StepController.h (ASLA
:167-197, the curvature controller :202-227, the no-progress stop :100-131;
no inertia veto, as the reference's default).
"""
from __future__ import annotations

import math

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
    """ColPivHouseholderQR of J^T (ContinuationSolver.h:119-136) over the DOFs
    that J touches (the other rows of J^T are zero: the same R)."""
    if J.shape[0] == 0:
        return np.zeros(0, np.int64)
    Jc = sp.csc_matrix(J)
    touched = np.nonzero(np.diff(Jc.indptr))[0]
    R, perm = sla.qr(Jc[:, touched].toarray().T, mode="r", pivoting=True)
    d = np.abs(np.diag(R))
    tol = max(J.shape) * np.finfo(float).eps * d.max()
    return np.sort(perm[:d.size][d > tol])


def _projector(Jd):
    """Returns (corrector(c), project(R)) for J_ind (dense or sparse); the Gram
    matrix J J^T (rows x rows) is solved dense."""
    if Jd.shape[0] == 0:
        return (lambda c, ns: np.zeros(ns)), (lambda R: R)
    Js = sp.csr_matrix(Jd)
    G = (Js @ Js.T).toarray()
    return (lambda c, ns: Js.T @ np.linalg.solve(G, -c)), (lambda R: R - Js.T @ np.linalg.solve(G, Js @ R))


def _kkt_spsolve(Ks, Js, rhs):
    A = sp.bmat([[Ks, Js.T], [Js, None]], format="csc") if Js.shape[0] else Ks.tocsc()
    return spla.spsolve(A, rhs)


def line_search(prob, lam, d, tol_force=1e-8, tol_energy=1e-8, tol_geom=1e-12, div=1e7, max_iters=500,
                kkt=None):
    """Returns (converged, iterations); d is updated in place (rolled back on failure).
    kkt(K_upper, J_ind, rhs) -> solution replaces the SuperLU solve of the
    assembled KKT matrix (large sheets)."""
    ns = d.size
    d0 = d.copy()
    c, J = prob.constraints(d)
    ind = _independent_rows(J)
    energy = 1.0
    for j in range(1, max_iters + 1):
        c, J = prob.constraints(d)
        ci, Jd = c[ind], sp.csr_matrix(J)[ind]
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
        rhs = np.concatenate([-(R + Ks @ xb), np.zeros(ci.size)])
        dx = (kkt(K, Jd, rhs) if kkt is not None else _kkt_spsolve(Ks, Jd, rhs))[:ns]
        alpha = 1.0
        for _ in range(6):
            dt = d + xb + alpha * dx
            Rt = prob.internal_force(dt) - lam * prob.external_force(dt)
            ct, Jt = prob.constraints(dt)
            _, pt = _projector(sp.csr_matrix(Jt)[ind])
            if np.abs(pt(Rt)).max() <= 1.05 * mf:
                break
            alpha *= 0.5
        energy = abs(np.dot(xb + alpha * dx, R))
        d[:] = dt
    d[:] = d0
    return False, -1


def load_path(prob, increments=20, lam_total=1.0, x0=None, first=1, on_increment=None, **tol):
    """Increments first..increments (lambda_k = k / increments * lam_total) from
    x0 (default 0); on_increment(k, x, iterations) after each."""
    x = np.zeros(prob.num_dofs) if x0 is None else np.array(x0, dtype=float)
    conv, its = [], []
    for k in range(first, increments + 1):
        ok, it = line_search(prob, lam_total * k / increments, x, **tol)
        conv.append(ok); its.append(it)
        if on_increment is not None:
            on_increment(k, x, it)
        if not ok:
            break
    return x, conv, its


def _kkt_solve(K, Jd, rhs):
    Ks = _sym(K)
    if Jd.shape[0]:
        A = sp.bmat([[Ks, sp.csr_matrix(Jd).T], [sp.csr_matrix(Jd), None]], format="csc")
    else:
        A = Ks.tocsc()
    return spla.splu(A).solve(rhs)


def arc_length(prob, lambda_0, x0, lambda_tot, on_converged=None, ds=0.1, dsmin=1e-6, dsmax=10.0, psi=1.0,
               kmax=5000, max_iters=100, Id=6, tol_force=1e-8, tol_energy=1e-8, tol_geom=1e-12, div=1e7):
    """Returns (x, lambda, [(step, lambda, corrector iterations)])."""
    ns = prob.num_dofs
    x = np.array(x0, dtype=float)
    lam, k = float(lambda_0), 1
    dU_old, dl_old = np.zeros(ns), 0.0
    c, J = prob.constraints(x)
    ind = _independent_rows(J)
    nc = ind.size
    steps = []
    np_ref, np_count, np_path = None, 0, 0.0

    def tangent(xx, ll):
        K, F = prob.assemble(xx)
        Kl = prob.load_stiffness(xx)
        return (K if Kl is None else K - ll * Kl), F

    while k <= kmax and lam <= lambda_tot:
        Fe = prob.external_force(x)
        K, _ = tangent(x, lam)
        c, J = prob.constraints(x)
        Jd = J[ind].toarray()
        corr, _ = _projector(Jd)
        xbp = corr(c[ind], ns)
        dT = _kkt_solve(K, Jd, np.concatenate([Fe, np.zeros(nc)]))[:ns]
        ff = Fe @ Fe
        if k == 1:
            sgn = 1.0 if dT @ Fe >= 0 else -1.0
        else:
            sgn = 1.0 if dU_old @ dT >= 0 else -1.0
        dl = sgn * ds / math.sqrt(dT @ dT + psi * psi * ff)
        dU = dl * dT + xbp
        ok, kc = False, max_iters
        for j in range(1, max_iters + 1):
            lc, xc = lam + dl, x + dU
            c, J = prob.constraints(xc)
            ci, Jd = c[ind], J[ind].toarray()
            Fec = prob.external_force(xc)
            Kc, Fi = tangent(xc, lc)
            R = Fi - lc * Fec
            vres = dU @ dU + psi * psi * dl * dl * ff - ds * ds
            corr, _ = _projector(Jd)
            xbc = corr(ci, ns)
            rhs = np.zeros((ns + nc, 2))
            rhs[:ns, 0] = Fec
            rhs[:ns, 1] = -(R + _sym(Kc) @ xbc)
            sol = _kkt_solve(Kc, Jd, rhs)
            daP, daPsi, mu = sol[:ns, 0], sol[:ns, 1] + xbc, sol[ns:, 1]
            dlc = -(0.5 * vres + dU @ daPsi) / (dU @ daP + psi * psi * dl * ff)
            dl += dlc
            dU = dU + daPsi + dlc * daP
            mf = np.abs(R + Jd.T @ mu).max()
            ge = np.abs(ci).max() if ci.size else 0.0
            if j > 1 and mf <= tol_force and abs(daPsi @ R) <= tol_energy and ge <= tol_geom and abs(vres) <= tol_geom:
                ok, kc = True, j
                break
            if mf >= div or np.isnan(mf):
                ok, kc = False, j
                break
        cth = 1.0
        if k > 1 and ok:
            a, b = psi * dl_old, psi * dl
            no, nw = math.sqrt(dU_old @ dU_old + a * a), math.sqrt(dU @ dU + b * b)
            if no > 0 and nw > 0:
                cth = min(1.0, max(-1.0, (dU_old @ dU + a * b) / (no * nw)))
        # ASLA
        opt = float(Id)
        if cth >= 0.99:
            opt = max(opt, 12.0)
        elif cth >= 0.90:
            opt = max(opt, 8.0)
        if not ok:
            ds *= 0.50 if kc < max_iters else 0.75
        else:
            sc = min(math.sqrt(opt / max(kc, 1)), 4.0)
            if 0.90 < sc < 1.10:
                sc = 1.0
            ds = min(max(ds * sc, dsmin), dsmax)
        if ok:
            if k > 1:       # curvature controller
                if cth >= 0.99:
                    ds = min(ds * 1.5, dsmax)
                elif cth >= 0.90:
                    pass
                elif cth >= 0.75:
                    ds = max(ds * 0.8, dsmin)
                else:
                    ds = max(ds * 0.5, dsmin)
            dU_old, dl_old = dU, dl
            x = x + dU
            lam += dl
            steps.append((k, lam, kc))
            if on_converged is not None:
                on_converged(k, lam, x)
            k += 1
            # no-progress stop at ds_min
            if ds <= dsmin * (1.0 + 1e-12):
                if np_ref is None:
                    np_ref, np_count, np_path = x.copy(), 0, 0.0
                else:
                    np_path += np.linalg.norm(dU)
                    np_count += 1
                    if np_count >= 24:
                        moved = np.linalg.norm(x - np_ref)
                        stalled = np_path > 0 and moved <= 0.15 * np_path
                        np_ref, np_count, np_path = x.copy(), 0, 0.0
                        if stalled:
                            break
            else:
                np_ref, np_count, np_path = None, 0, 0.0
        elif ds <= dsmin:
            break
    return x, lam, steps
