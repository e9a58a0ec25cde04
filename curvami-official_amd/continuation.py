"""Constrained Newton with backtracking and the load-increment path.

The host side of the Newton loop that calls the device assembly, with the
semantics of ContinuationSolver::LineSearch
(core/include/curvami/solver/ContinuationSolver.h:185-322):

* constraint rows are filtered once per call to an independent set by a
  column-pivoted QR of J^T with the tolerance max(rows, cols) * eps * max|R_ii|
  (InitializeTopology, :105-137) and reused (filterRows, :139-164);
* the force residual R = F_int - lambda F_ext is projected onto the null space
  of J_ind (ConstraintProjector, core/include/curvami/solver/ConstraintProjector.h:19-55);
* convergence at iteration j > 1 when max|P R| <= tol_force, |dx . R| <= tol_energy
  and max|c| <= tol_geom (:225-233); divergence when max|P R| >= div or NaN (:235-239);
* the step is the geometric corrector xb = -J^T (J J^T)^-1 c plus the KKT step
  [K J^T; J 0] [dx; mu] = [-(R + K xb); 0] (:241-263), K read as the upper triangle;
* backtracking: alpha = 1, 1/2, ... (6 trials) until the projected residual of the
  trial is <= 1.05 times the current one; without descent the last trial is kept
  unless fail_on_no_descent (:268-312);
* on failure the state rolls back to the call's start (:238, :305, :318-321).

Arc-length continuation (ContinuationSolver::ArcLength, :329-538) with the
step control of core/include/curvami/solver/StepController.h: Crisfield's
constraint |dU|^2 + psi^2 dlambda^2 |F_ext|^2 = ds^2, a tangent predictor whose
sign follows the previous increment (snap-back reversal), corrector iterations
with two KKT right-hand sides, ASLA and the curvature controller for ds, the
no-progress stop at ds_min, the accumulated external work, and the restart
record (RestartManager, read and written through the library's `.res` codec).
The inertia veto of a reversal (snapback_inertia_veto, StepController.h:25-44,
off by default as in the reference) reads the KKT factorization's negative
pivot count.

The KKT solve replaces the reference's symmetric MUMPS factorization
(core/include/curvami/solver/MumpsSymmetricSolver.h:46-55, :150-230) by the
library's host solver (libcurvami_kkt: a sparse symmetric indefinite LDL^T whose
symbolic analysis is kept while the pattern is unchanged, with the negative
pivot count of every factorization, MUMPS' INFOG(12)); SuperLUKKTSolver (scipy)
remains for cross-checks. The constraint algebra is sparse: the independent
rows come from a column-pivoted QR of J^T restricted to the DOFs J touches
(the other rows of J^T are zero, so R is that of the full QR), and the
projector factors the sparse Gram matrix J J^T.

With an inexact (FP32) tangent the iteration is safeguarded: an FP32-tangent
step that reduces the merit by less than fp32_stall_ratio switches the rest of
the increment to the FP64 tangent, and after fp32_stall_limit such increments
the solver keeps the FP64 tangent (on large sheets folding is nearly rigid for
the facets, so the face tangent's action along the fold cancels down to the
hinge's small share, below the FP32 rounding of the face blocks, and the
inexact iteration stalls); finish_fp64 ends each LineSearch with one Newton
step on the FP64 tangent once the stopping rule holds, so that the converged
state is that of the quadratically converging FP64 iteration.

A problem object provides the Problem interface
(core/include/curvami/solver/Problem.h:16-68): num_dofs, assemble(x) -> (K upper
as scipy CSC, F_int), internal_force(x), external_force(x), constraints(x) ->
(c, J as scipy CSR), load_stiffness(x) -> K_load upper or None.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import List, Optional

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def full_from_upper(K):
    """upper(K) + strictly-upper(K)^T: the selfadjointView<Upper> of a CSC
    matrix. The assembled K stores its diagonal node blocks as full 3 x 3, so
    entries below the diagonal exist and must be ignored, as MUMPS (sym = 2)
    ignores them (MumpsSymmetricSolver.h:84-126)."""
    return (sp.triu(K) + sp.triu(K, 1).T).tocsc()


_LOWER_CACHE = {}


def sym_matvec(K, v):
    """selfadjointView<Upper>(K) v without forming the symmetric matrix:
    U v + U^T v - diag(U) v - L v - L^T v, with L the stored entries below the
    diagonal (their positions cached per pattern)."""
    K = K.tocsc()
    key = (id(K.indptr), id(K.indices), K.shape, K.nnz)
    hit = _LOWER_CACHE.get(key)
    if hit is None or hit[0] is not K.indptr:
        col = np.repeat(np.arange(K.shape[1], dtype=np.int64), np.diff(K.indptr))
        low = np.nonzero(K.indices > col)[0]
        hit = (K.indptr, low, K.indices[low].astype(np.int64), col[low])
        _LOWER_CACHE.clear()
        _LOWER_CACHE[key] = hit
    _, low, lr, lc = hit
    y = K @ v + K.T @ v - K.diagonal() * v
    if low.size:
        lv = K.data[low]
        y -= np.bincount(lr, lv * v[lc], minlength=y.size) + np.bincount(lc, lv * v[lr], minlength=y.size)
    return y


class ConstraintProjector:
    """(J J^T) factorized once, used for the minimum-norm corrector and the
    null-space projection (ConstraintProjector.h:19-55). The Gram matrix is
    formed sparse (rows couple only through shared DOFs) and factored by a
    sparse LU (a dense Cholesky for up to DENSE_ROWS rows); a failed
    factorization falls back to a least-squares solve."""

    DENSE_ROWS = 512

    def __init__(self, J):
        self.J = J.tocsr()
        self.nr = J.shape[0]
        self._chol = None
        self._lu = None
        if self.nr:
            G = (self.J @ self.J.T).tocsc()
            self._G = G
            try:
                if self.nr <= self.DENSE_ROWS:
                    self._chol = sla.cho_factor(G.toarray(), lower=True, check_finite=True)
                else:
                    self._lu = spla.splu(G)
            except (np.linalg.LinAlgError, ValueError, RuntimeError):
                self._chol = self._lu = None

    def _gram_solve(self, b):
        if self._chol is not None:
            y = sla.cho_solve(self._chol, b)
            if np.all(np.isfinite(y)):
                return y
        if self._lu is not None:
            y = self._lu.solve(b)
            if np.all(np.isfinite(y)):
                return y
        return spla.lsqr(self._G, b, atol=1e-15, btol=1e-15)[0]

    def corrector(self, c, ns):
        if c.size == 0:
            return np.zeros(ns)
        return self.J.T @ self._gram_solve(-c)

    def project(self, R):
        if self.nr == 0:
            return R
        return R - self.J.T @ self._gram_solve(self.J @ R)


class KKTSolver:
    """[K J^T; J 0] (K given as its stored upper triangle) by the library's host
    LDL^T (MumpsSymmetricSolver's role, MumpsSymmetricSolver.h:150-230): the
    symbolic analysis is redone only when the pattern of K or J changes;
    neg_pivots is the negative pivot count of the last factorization
    (INFOG(12), :58-60, :217)."""

    def __init__(self, nthreads: int = 0):
        import curvami_gpu as G
        self._s = G.KKT(nthreads)
        self._err = G.CvmError
        self.factorizations = 0
        self.last_error = 0
        self.neg_pivots = -1

    def solve(self, K_upper, J, rhs):
        K = K_upper.tocsc()
        J = J.tocsr()
        if not K.has_sorted_indices:
            K = K.sorted_indices()
        rows = np.repeat(np.arange(J.shape[0], dtype=np.int32), np.diff(J.indptr))
        self.last_error = 0
        try:
            self._s.set_pattern(K.indptr, K.indices, J.shape[0], rows, J.indices)
            self.neg_pivots, _ = self._s.factorize(K.data, J.data)
            self.factorizations += 1
            sol = self._s.solve(rhs)
        except self._err:
            self.last_error = -1
            return None
        if not np.all(np.isfinite(sol)):
            self.last_error = -2
            return None
        return sol

    def stats(self):
        return self._s.stats()


class SuperLUKKTSolver:
    """[K J^T; J 0] by scipy's SuperLU on the assembled symmetric matrix, for
    cross-checks (ordering and factorization on every call; no inertia)."""

    def __init__(self, permc_spec: str = "MMD_AT_PLUS_A"):
        self.permc_spec = permc_spec
        self.factorizations = 0
        self.last_error = 0
        self.neg_pivots = -1

    def solve(self, K_upper, J, rhs):
        ns = K_upper.shape[0]
        Kf = full_from_upper(K_upper)
        if J.shape[0]:
            A = sp.bmat([[Kf, J.T], [J, None]], format="csc")
        else:
            A = Kf
        try:
            lu = spla.splu(A, permc_spec=self.permc_spec)
        except RuntimeError:
            self.last_error = -1
            return None
        self.factorizations += 1
        sol = lu.solve(rhs)
        if not np.all(np.isfinite(sol)):
            self.last_error = -2
            return None
        return sol


@dataclasses.dataclass
class ArcLengthStep:
    step: int
    lam: float
    iterations: int            # corrector iterations of the accepted step
    ds: float                  # arc length after the step control
    W_ext: float               # accumulated external work


@dataclasses.dataclass
class ArcLengthReport:
    x: np.ndarray
    lam: float
    steps: List[ArcLengthStep]
    stop: str                  # "lambda_tot", "kmax", "ds_min", "kkt", "stalled"
    reversals: int = 0
    restarted_from: int = 0    # step loaded from the restart record (0: none)
    vetoed: int = 0            # reversals vetoed by the unchanged inertia


@dataclasses.dataclass
class LineSearchReport:
    converged: bool
    iterations: int            # Newton iterations at convergence (-1 otherwise)
    attempts: int
    history: List[tuple]       # (j, trial merit, energy change, geometric error)
    fp64_steps: int = 0        # FP64 tangents used by an inexact-tangent LineSearch (finishing step, stall)
    fp32_stalled: bool = False  # the inexact iteration stalled and switched to the FP64 tangent


class ContinuationSolver:
    tol_force = 1e-8
    tol_energy = 1e-8
    tol_geom = 1e-12
    div = 1e7
    ls_max_iters = 500
    fail_on_no_descent = False
    # arc-length continuation and step control (ContinuationSolver.h:31-44,
    # StepController.h:16-24, :51-80)
    kmax = 5000
    max_iters = 100
    ds = 0.1
    dsmin = 1e-6
    dsmax = 10.0
    Id = 6
    psi = 1.0
    restart_file = ""
    asla_hold_after_reversal = False
    snapback_inertia_veto = False
    finish_fp64 = True          # inexact (FP32) tangents: one FP64 Newton step after the stopping rule holds
    fp32_stall_ratio = 0.9      # inexact tangents: a step keeping more of the merit switches to FP64
    fp32_stall_limit = 2        # after this many stalled increments the FP32 tangent is given up
    no_progress_window = 24
    no_progress_efficiency = 0.15
    no_progress_abort = True
    NO_REVERSAL_YET = 1 << 20

    def __init__(self, kkt: Optional[KKTSolver] = None, verbose: bool = False):
        self.kkt = kkt or KKTSolver()
        self.ind_rows = np.zeros(0, np.int64)
        self._prev_neg_modes = -1
        self.fp32_stalls = 0         # increments whose FP32-tangent iteration stalled
        self.verbose = verbose
        self.last_report: Optional[LineSearchReport] = None

    # ---- InitializeTopology / filterRows (ContinuationSolver.h:105-164) -------
    def initialize_topology(self, J):
        """ColPivHouseholderQR of J^T with the threshold max(rows, cols) eps
        max|R_ii| (:119-136), on the DOFs J touches only: the other rows of J^T
        are zero and leave R unchanged."""
        if J.shape[0] == 0 or J.shape[1] == 0:
            self.ind_rows = np.zeros(0, np.int64)
            return
        J = J.tocsc()
        if not np.all(np.isfinite(J.data)):
            raise RuntimeError("Constraint Jacobian contains NaN or Inf.")
        touched = np.nonzero(np.diff(J.indptr))[0]
        Jt = J[:, touched].toarray().T
        R, perm = sla.qr(Jt, mode="r", pivoting=True)
        d = np.abs(np.diag(R))
        tol = max(J.shape) * np.finfo(np.float64).eps * (d.max() if d.size else 0.0)
        self.ind_rows = np.sort(perm[:d.size][d > tol])

    def filter_rows(self, c, J):
        return c[self.ind_rows], J.tocsr()[self.ind_rows]

    # ---- step control (StepController.h) ---------------------------------------
    def _reset_tracking(self):
        self._since_reversal = self.NO_REVERSAL_YET
        self._prev_neg_modes = -1
        self._np_ref = None
        self._np_count = 0
        self._np_path = 0.0

    def stalled_at_dsmin(self, x, dU) -> bool:
        """Net movement over a window of accepted steps at ds_min against the path
        length walked (StepController.h:100-131)."""
        if not self.no_progress_abort:
            return False
        if self.ds > self.dsmin * (1.0 + 1e-12):
            self._np_ref, self._np_count, self._np_path = None, 0, 0.0
            return False
        if self._np_ref is None or self._np_ref.size != x.size:
            self._np_ref, self._np_count, self._np_path = x.copy(), 0, 0.0
            return False
        self._np_path += float(np.linalg.norm(dU))
        self._np_count += 1
        if self._np_count < self.no_progress_window:
            return False
        moved = float(np.linalg.norm(x - self._np_ref))
        path = self._np_path
        self._np_ref, self._np_count, self._np_path = x.copy(), 0, 0.0
        return path > 0.0 and moved <= self.no_progress_efficiency * path

    def asla(self, ok: bool, k_corr: int, cos_theta: float):
        """Crisfield's square-root rule with target relaxation on straight paths
        (StepController.h:167-197)."""
        opt = float(self.Id)
        if cos_theta >= 0.99:
            opt = max(opt, 12.0)
        elif cos_theta >= 0.90:
            opt = max(opt, 8.0)
        if not ok:
            self.ds *= 0.50 if k_corr < self.max_iters else 0.75
            return
        scale = min(math.sqrt(opt / max(k_corr, 1)), 4.0)
        if 0.90 < scale < 1.10:
            scale = 1.0
        self.ds = min(max(self.ds * scale, self.dsmin), self.dsmax)

    def curvature_controller(self, step: int, cos_theta: float):
        """Tiered ds control on the angle between successive increments
        (StepController.h:202-227)."""
        if step <= 1:
            return
        if cos_theta >= 0.99:
            if not (self.asla_hold_after_reversal and self._since_reversal <= 1):
                self.ds = min(self.ds * 1.5, self.dsmax)
        elif cos_theta >= 0.90:
            pass
        elif cos_theta >= 0.75:
            self.ds = max(self.ds * 0.8, self.dsmin)
        else:
            self.ds = max(self.ds * 0.5, self.dsmin)

    # ---- inertia tracking (StepController.h:25-44, :125-160) --------------------
    def _neg_modes(self, n_const):
        """Negative modes of the reduced Hessian: negative KKT pivots of the last
        factorization minus the constraint rows (ContinuationSolver.h:69-78)."""
        neg = getattr(self.kkt, "neg_pivots", -1)
        return neg - n_const if neg >= 0 else -1

    def confirm_reversal(self, neg_modes) -> bool:
        """A requested reversal goes ahead unless the veto is on, a baseline
        exists and the inertia did not change; the tracker always advances."""
        prev, self._prev_neg_modes = self._prev_neg_modes, neg_modes
        if not self.snapback_inertia_veto or prev < 0 or neg_modes < 0:
            return True
        return neg_modes != prev

    def _apply_load_stiffness(self, prob, K, x, lam):
        Kl = prob.load_stiffness(x)
        return K if Kl is None else (K - lam * Kl).tocsc()

    # ---- ArcLength (ContinuationSolver.h:329-538) --------------------------------
    def arc_length(self, prob, lambda_0: float, x0, lambda_tot: float, on_converged=None) -> ArcLengthReport:
        """Continuation from (x0, lambda_0) while lambda <= lambda_tot and at most
        kmax steps; on_converged(step, lambda, x, W_ext) after every accepted
        step. The solver's ds is adapted in place, as in the reference."""
        ns = prob.num_dofs
        x = np.array(x0, dtype=np.float64)
        if x.size != ns:
            raise ValueError(f"ArcLength: state vector has {x.size} entries, problem has {ns} DOFs")
        lam = float(lambda_0)
        k = 1
        dU_old = np.zeros(ns)
        dl_old = 0.0
        c_all, J_all = prob.constraints(x)
        self.initialize_topology(J_all)
        self._reset_tracking()
        n_const = int(self.ind_rows.size)
        W_ext = 0.0
        F_ref = np.array(prob.external_force(x), dtype=np.float64)
        x_prev = x.copy()
        rep = ArcLengthReport(x, lam, [], "lambda_tot")
        if self.restart_file and os.path.exists(self.restart_file):
            import curvami_gpu
            r = curvami_gpu.restart_load(self.restart_file, ns)
            k, lam, self.ds, dl_old, W_ext = r["k"], r["lam"], r["ds"], r["dl_old"], r["W_ext"]
            x[:] = r["x"]
            dU_old = r["dU_old"]
            x_prev = x.copy()      # work increments are measured from the loaded state
            rep.restarted_from = k
        psi2 = self.psi * self.psi
        while k <= self.kmax and lam <= lambda_tot:
            # ---- predictor: tangent d_T = K^-1 F_ext under the constraints
            Fext = prob.external_force(x)
            K, _ = prob.assemble(x)
            K = self._apply_load_stiffness(prob, K, x, lam)
            c_all, J_all = prob.constraints(x)
            c_ind, J_ind = self.filter_rows(c_all, J_all)
            xb_p = ConstraintProjector(J_ind).corrector(c_ind, ns)
            rhs = np.zeros(ns + n_const)
            rhs[:ns] = Fext
            sol = self.kkt.solve(K, J_ind, rhs)
            if sol is None:
                rep.stop = "kkt"
                break
            d_T = sol[:ns]
            ff = float(np.dot(Fext, Fext))
            if k == 1:
                sgn = 1.0 if float(np.dot(d_T, Fext)) >= 0.0 else -1.0
            elif float(np.dot(dU_old, d_T)) >= 0.0:
                sgn = 1.0
                self._prev_neg_modes = self._neg_modes(n_const)
            elif self.confirm_reversal(self._neg_modes(n_const)):
                # snap-back: the tangent turned against the last increment
                sgn = -1.0
                self._since_reversal = 0
                rep.reversals += 1
            else:
                sgn = 1.0
                rep.vetoed += 1
            dl = sgn * self.ds / math.sqrt(float(np.dot(d_T, d_T)) + psi2 * ff)
            dU = dl * d_T + xb_p
            ok, k_corr = False, self.max_iters
            # ---- corrector
            for j in range(1, self.max_iters + 1):
                lam_c = lam + dl
                x_c = x + dU
                c_all, J_all = prob.constraints(x_c)
                c_ind_c, J_ind_c = self.filter_rows(c_all, J_all)
                Fext_c = prob.external_force(x_c)
                K_c, Fint_c = prob.assemble(x_c)
                K_c = self._apply_load_stiffness(prob, K_c, x_c, lam_c)
                R = Fint_c - lam_c * Fext_c
                v_res = float(np.dot(dU, dU)) + psi2 * dl * dl * ff - self.ds * self.ds
                xb_c = ConstraintProjector(J_ind_c).corrector(c_ind_c, ns)
                rhs2 = np.zeros((ns + n_const, 2))
                rhs2[:ns, 0] = Fext_c
                rhs2[:ns, 1] = -(R + sym_matvec(K_c, xb_c))
                sol2 = self.kkt.solve(K_c, J_ind_c, rhs2)
                if sol2 is None:
                    ok, k_corr = False, j
                    break
                daP = sol2[:ns, 0]
                daPsi = sol2[:ns, 1] + xb_c
                mult = sol2[ns:, 1]
                dl_c = -(0.5 * v_res + float(np.dot(dU, daPsi))) / (float(np.dot(dU, daP)) + psi2 * dl * ff)
                dl += dl_c
                dU = dU + daPsi + dl_c * daP
                max_force = float(np.abs(R + J_ind_c.T @ mult).max()) if ns else 0.0
                energy_change = abs(float(np.dot(daPsi, R)))
                geom_error = float(np.abs(c_ind_c).max()) if c_ind_c.size else 0.0
                if j > 1 and max_force <= self.tol_force and energy_change <= self.tol_energy \
                        and geom_error <= self.tol_geom and abs(v_res) <= self.tol_geom:
                    ok, k_corr = True, j
                    break
                if max_force >= self.div or math.isnan(max_force):
                    ok, k_corr = False, j
                    break
                if self.verbose:
                    print(f"{j:3d}  {max_force:12.5e}  {energy_change:12.5e}  {geom_error:12.5e}  {abs(v_res):12.5e}",
                          flush=True)
            # ---- step control
            cos_theta = 1.0
            if k > 1 and ok:
                old_load, new_load = self.psi * dl_old, self.psi * dl
                old_n = math.sqrt(float(np.dot(dU_old, dU_old)) + old_load * old_load)
                new_n = math.sqrt(float(np.dot(dU, dU)) + new_load * new_load)
                if old_n > 0.0 and new_n > 0.0:
                    cos_theta = (float(np.dot(dU_old, dU)) + old_load * new_load) / (old_n * new_n)
                    cos_theta = max(-1.0, min(1.0, cos_theta))
            self.asla(ok, k_corr, cos_theta)
            if ok:
                self.curvature_controller(k, cos_theta)
                if self._since_reversal < self.NO_REVERSAL_YET:
                    self._since_reversal += 1
                dU_old, dl_old = dU, dl
                x += dU
                lam += dl
                W_ext += 0.5 * ((lam - dl) + lam) * float(np.dot(F_ref, x - x_prev))
                x_prev = x.copy()
                rep.steps.append(ArcLengthStep(k, lam, k_corr, self.ds, W_ext))
                if on_converged is not None:
                    on_converged(k, lam, x, W_ext)
                k += 1
                if self.restart_file:
                    import curvami_gpu
                    curvami_gpu.restart_save(self.restart_file, k, lam, self.ds, dl_old, W_ext, x, dU_old)
                if self.stalled_at_dsmin(x, dU):
                    rep.stop = "stalled"
                    break
            elif self.ds <= self.dsmin:
                rep.stop = "ds_min"
                break
        else:
            rep.stop = "kmax" if k > self.kmax else "lambda_tot"
        rep.x, rep.lam = x, lam
        return rep

    # ---- LineSearch (ContinuationSolver.h:185-322) -----------------------------
    def line_search(self, prob, lam: float, d: np.ndarray) -> LineSearchReport:
        ns = d.size
        if ns != prob.num_dofs:
            raise ValueError(f"LineSearch: state vector has {ns} entries, problem has {prob.num_dofs} DOFs")
        d0 = d.copy()
        c_all, J_all = prob.constraints(d)
        self.initialize_topology(J_all)
        energy_change = 1.0
        hist = []
        rep = LineSearchReport(False, -1, 0, hist)
        self.last_report = rep
        inexact = getattr(prob, "fp32_tangent", False) and hasattr(prob, "assemble_fp64")
        finish = self.finish_fp64 and inexact
        fp64_step = False           # the next tangent is FP64 (finishing step)
        # every further tangent of this increment is FP64: after a stall, or from
        # the start once the FP32 tangent has stalled fp32_stall_limit times
        fp64_rest = inexact and self.fp32_stalls >= self.fp32_stall_limit
        for j in range(1, self.ls_max_iters + 1):
            rep.attempts = j
            c_all, J_all = prob.constraints(d)
            c_ind, J_ind = self.filter_rows(c_all, J_all)
            R = prob.internal_force(d) - lam * prob.external_force(d)
            proj = ConstraintProjector(J_ind)
            max_force = float(np.abs(proj.project(R)).max()) if ns else 0.0
            geom_error = float(np.abs(c_ind).max()) if c_ind.size else 0.0
            if j > 1 and max_force <= self.tol_force and energy_change <= self.tol_energy \
                    and geom_error <= self.tol_geom:
                if finish and not fp64_step and not fp64_rest:
                    fp64_step = True       # one more step, on the FP64 tangent
                else:
                    rep.converged = True
                    rep.iterations = j
                    return rep
            if max_force >= self.div or math.isnan(max_force):
                d[:] = d0
                return rep
            use64 = fp64_step or fp64_rest
            K, _ = prob.assemble_fp64(d) if use64 else prob.assemble(d)
            if use64 and inexact:
                rep.fp64_steps += 1
            Kl = prob.load_stiffness(d)
            if Kl is not None:
                K = (K - lam * Kl).tocsc()
            xb = proj.corrector(c_ind, ns)
            rhs = np.zeros(ns + c_ind.size)
            rhs[:ns] = -(R + sym_matvec(K, xb))
            sol = self.kkt.solve(K, J_ind, rhs)
            if sol is None:
                d[:] = d0
                return rep
            dx = sol[:ns]
            alpha = 1.0
            trial = max_force
            descended = False
            d_trial = d
            for _ in range(6):
                d_trial = d + xb + alpha * dx
                R_t = prob.internal_force(d_trial) - lam * prob.external_force(d_trial)
                c_t, J_t = prob.constraints(d_trial)
                _, J_t_ind = self.filter_rows(c_t, J_t)
                trial = float(np.abs(ConstraintProjector(J_t_ind).project(R_t)).max())
                if trial <= 1.05 * max_force:
                    descended = True
                    break
                alpha *= 0.5
            if not descended and self.fail_on_no_descent:
                d[:] = d0
                return rep
            energy_change = abs(float(np.dot(xb + alpha * dx, R)))
            if inexact and not use64 and trial > self.fp32_stall_ratio * max_force:
                fp64_rest = True
                rep.fp32_stalled = True
                self.fp32_stalls += 1
            d[:] = d_trial
            hist.append((j, trial, energy_change, geom_error))
            if self.verbose:
                print(f"{j:3d}  {trial:12.5e}  {energy_change:12.5e}  {geom_error:12.5e}", flush=True)
        d[:] = d0
        return rep


def initial_condition_with_folds(prob_rigid, solver: "ContinuationSolver") -> np.ndarray:
    """NonLinearSolver::GetInitialConditionWithFolds (plate/NonLinearSolver.h:101-148):
    from x = 0, the minimum-norm corrector xb of every constraint row (fold
    angle rows included: prob_rigid is built with flag_rig = True) plus the KKT
    solution of [K J^T; J 0] [dx; mu] = [-K xb; 0]."""
    n = prob_rigid.num_dofs
    x = np.zeros(n)
    c, J = prob_rigid.constraints(x)
    if c.size == 0:
        return x
    solver.initialize_topology(J)
    c_ind, J_ind = solver.filter_rows(c, J)
    K, _ = prob_rigid.assemble(x)
    xb = ConstraintProjector(J_ind).corrector(c_ind, n)
    rhs = np.zeros(n + c_ind.size)
    rhs[:n] = -sym_matvec(K, xb)
    sol = solver.kkt.solve(K, J_ind, rhs)
    if sol is None:
        raise RuntimeError(f"initial folded-condition KKT solve failed with error code {solver.kkt.last_error}")
    return xb + sol[:n]


@dataclasses.dataclass
class LoadPathResult:
    x: np.ndarray
    converged: List[bool]
    iterations: List[int]
    lambdas: List[float]
    fp64_steps: List[int] = dataclasses.field(default_factory=list)   # finishing steps per increment

    @property
    def all_converged(self) -> bool:
        return all(self.converged)


def load_path(prob, increments: int = 20, lam_total: float = 1.0, x0=None,
              solver: Optional[ContinuationSolver] = None, stop_on_failure: bool = True,
              on_increment=None) -> LoadPathResult:
    """Proportional loading lambda_k = k / increments * lam_total, one LineSearch
    per increment from the previous converged state; on_increment(k, x, report)
    after each."""
    solver = solver or ContinuationSolver()
    x = np.zeros(prob.num_dofs) if x0 is None else np.array(x0, dtype=np.float64)
    conv, its, lams, f64 = [], [], [], []
    for k in range(1, increments + 1):
        lam = lam_total * k / increments
        rep = solver.line_search(prob, lam, x)
        conv.append(rep.converged); its.append(rep.iterations); lams.append(lam); f64.append(rep.fp64_steps)
        if on_increment is not None:
            on_increment(k, x, rep)
        if not rep.converged and stop_on_failure:
            break
    return LoadPathResult(x, conv, its, lams, f64)


class PlateProblem:
    """PlateProblem (plate/PlateProblem.h:52-91) over the device plate: residual,
    tangent, constraint rows and external load all come from the HIP library.
    fp32_tangent assembles the face tangent in FP32 (inexact Newton);
    exact_hinge adds the hinge's (theta - theta0) d2theta term (fold_exact_tangent,
    plate/Meshless.h:48-50)."""

    def __init__(self, plate, flag_rig: bool = False, fp32_tangent: bool = False, exact_hinge: bool = False):
        self.plate = plate
        self.flag_rig = flag_rig
        self.fp32_tangent = fp32_tangent
        self.exact_hinge = exact_hinge
        self.num_dofs = 3 * plate.nn
        outer, inner = plate.pattern(with_folds=not flag_rig)
        self._outer, self._inner = outer, inner
        self._fext = None
        self._rows, self._cols = plate.constraint_layout(rigid=flag_rig)
        self.assemblies = 0
        self.residuals = 0

    def assemble(self, x, fp32_tangent=None):
        fp32 = self.fp32_tangent if fp32_tangent is None else fp32_tangent
        F, K = self.plate.assemble(x, elastic_folds=not self.flag_rig, fp32_tangent=fp32,
                                   exact_hinge=self.exact_hinge)
        self.assemblies += 1
        n = self.num_dofs
        return sp.csc_matrix((K, self._inner, self._outer), shape=(n, n)), F

    def assemble_fp64(self, x):
        """The tangent in FP64 whatever fp32_tangent says (the finishing step)."""
        return self.assemble(x, fp32_tangent=False)

    def internal_force(self, x):
        F, _ = self.plate.assemble(x, elastic_folds=not self.flag_rig, tangent=False)
        self.residuals += 1
        return F

    def external_force(self, x):
        if self._fext is None:
            self._fext = self.plate.external_force()
        return self._fext

    def load_stiffness(self, x):
        return None

    def constraints(self, x):
        c, vals = self.plate.constraints(x)
        J = sp.csr_matrix((vals, (self._rows, self._cols)), shape=(c.size, self.num_dofs))
        return c, J
