"""Progressive creasing: a crease written along a fixed path, its tips moving
from vertex to vertex, one Newton solve per step
(Progressive::run, plate/ProgressiveCreasing.h:384-872; the driver's set-up,
main.cpp:536-611).

Per step: the cut is the path truncated at the tips (end caps at tips inside
the sheet), every shape function is rebuilt with it (the device re-cut:
cvm_clear_creases + cvm_add_crease + cvm_shape_functions_all), crease GPs
between the tips are switched on with rest angle theta0* (hinge and tie rows
only there: cvm_set_fold_gp_state), and one LineSearch at lambda = 0 runs from
the carried-over coefficients. A failed (or guard-refused) step retries by
raising theta0 on the newest crease points over `fallback_substeps`
sub-steps. Energies (cvm_energies) and crease angles (cvm_fold_angles) are
device results; the per-step bookkeeping here is host logic.

Each accepted step can be exported through `on_step` (export.step_record /
StepWriter). The negative-mode count of each step is read from the last KKT
factorization of its Newton solve, as lastNegativeModes does
(ProgressiveCreasing.h:144-153; None where that factorization reported no
inertia). Not reproduced: the coefficient re-fit (refit other than "none")
and the diagnose_step mode.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable, List, Optional

import numpy as np

from workloads import Crease


# ---------------------------------------------------------------------------
# path geometry (ProgressiveCreasing.h:86-121)
# ---------------------------------------------------------------------------

def cumulative_arc_length(cr: Crease) -> np.ndarray:
    S = np.zeros(len(cr.vx))
    for i in range(1, S.size):
        S[i] = S[i - 1] + math.sqrt((cr.vx[i] - cr.vx[i - 1]) ** 2 + (cr.vy[i] - cr.vy[i - 1]) ** 2)
    return S


def project_arc_length(px: float, py: float, cr: Crease, S: np.ndarray) -> float:
    """Arc length of the point of the path closest to (px, py)."""
    best_d, best_s = math.inf, 0.0
    for i in range(len(cr.vx) - 1):
        ax, ay = cr.vx[i], cr.vy[i]
        vx, vy = cr.vx[i + 1] - ax, cr.vy[i + 1] - ay
        v2 = vx * vx + vy * vy
        if v2 == 0.0:
            continue
        t = max(0.0, min(1.0, ((px - ax) * vx + (py - ay) * vy) / v2))
        d = math.sqrt((px - (ax + t * vx)) ** 2 + (py - (ay + t * vy)) ** 2)
        if d < best_d:
            best_d, best_s = d, S[i] + t * math.sqrt(v2)
    return best_s


def truncate(full: Crease, a: int, b: int) -> Crease:
    """The cut between path vertices a and b: tips inside the sheet get the end
    cap (extend false); an end of the full path keeps its flag."""
    N = len(full.vx) - 1
    return Crease(np.array(full.vx[a:b + 1]), np.array(full.vy[a:b + 1]),
                  full.ext_s if a == 0 else False, full.ext_e if b == N else False,
                  bool(full.closed and a == 0 and b == N))


@dataclasses.dataclass
class Stage:
    a: int
    b: int
    frac: float
    outside: bool = False


def schedule(mode: str, tip_step: int, num_steps: int, N: int, skip=()) -> List[Stage]:
    """Tip positions per step (ProgressiveCreasing.h:396-441)."""
    st: List[Stage] = []
    m = max(1, tip_step)
    skipped = lambda v: v in skip
    if mode == "all_at_once":
        st = [Stage(0, N, k / num_steps) for k in range(num_steps + 1)]
    elif mode == "one_end":
        k = 0
        while True:
            b = min(k * m, N)
            if b == N or not skipped(b):
                st.append(Stage(0, b, 1.0))
            if b == N:
                break
            k += 1
    elif mode == "one_end_reverse":
        k = 0
        while True:
            a = max(N - k * m, 0)
            if a == 0 or not skipped(a):
                st.append(Stage(a, N, 1.0))
            if a == 0:
                break
            k += 1
    elif mode == "middle_out":
        c = N // 2
        k = 0
        while True:
            a, b = max(c - k * m, 0), min(c + k * m, N)
            st.append(Stage(a, b, 1.0))
            if a == 0 and b == N:
                break
            k += 1
    elif mode == "ends_in":
        c = N // 2
        k = 0
        while True:
            a, b = min(k * m, c), max(N - k * m, c)
            last = a == c and b == c
            if last or (not skipped(a) and not skipped(b)):
                st.append(Stage(a, b, 1.0, True))
            if last:
                break
            k += 1
    else:
        raise ValueError(f"Progressive creasing: unknown mode '{mode}'")
    return st


# ---------------------------------------------------------------------------
# fold-angle statistics (plate/FoldAngleUtil.h:26-60, ProgressiveCreasing.h:157-179)
# ---------------------------------------------------------------------------

def circular_mean(theta, w) -> float:
    s = float(np.sum(np.sin(theta) * w))
    c = float(np.sum(np.cos(theta) * w))
    return math.atan2(s, c)


def unwrap_about(theta, ref):
    d = np.asarray(theta, np.float64) - ref
    d = np.where(d > math.pi, d - 2 * math.pi * np.ceil((d - math.pi) / (2 * math.pi)), d)
    d = np.where(d < -math.pi, d + 2 * math.pi * np.ceil((-math.pi - d) / (2 * math.pi)), d)
    return ref + d


def active_fold_angles(theta, w, active):
    """Mean (about the circular mean) and largest |theta| over the active
    crease GPs, in degrees."""
    on = np.asarray(active, bool)
    th, ww = np.asarray(theta)[on], np.asarray(w)[on]
    if th.size == 0:
        return 0.0, 0.0
    ref = circular_mean(th, ww)
    u = unwrap_about(th, ref)
    wsum = float(ww.sum())
    mean = float(np.sum(u * ww) / wsum) if wsum > 1e-14 else 0.0
    return mean * 180.0 / math.pi, float(np.abs(u).max()) * 180.0 / math.pi


# ---------------------------------------------------------------------------
# the making
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class CreasingConfig:
    """Progressive::Config (ProgressiveCreasing.h:51-84) and main.cpp:538-556."""
    num_steps: int = 40
    fallback_substeps: int = 4
    L_ref: float = 1.0
    guard_max_theta_deg: float = 180.0
    guard_energy_factor: float = 3.0
    guard_max_motion: float = 0.5
    guard_max_motion_topology: float = 1.0

    @classmethod
    def from_json(cls, pc: dict) -> "CreasingConfig":
        if pc.get("refit", "none") not in ("none", False):
            raise NotImplementedError("creasing.refit: only 'none' is supported")
        g = pc.get("guard", {})
        d = cls()
        return cls(int(pc.get("num_steps", 40)), int(pc.get("fallback_substeps", 4)), float(pc["L_ref"]),
                   float(g.get("max_theta_deg", d.guard_max_theta_deg)),
                   float(g.get("energy_factor", d.guard_energy_factor)),
                   float(g.get("max_motion", d.guard_max_motion)),
                   float(g.get("max_motion_topology", d.guard_max_motion_topology)))


@dataclasses.dataclass
class Line:
    folds: List[int]
    full: Crease
    mode: str = "one_end"
    tip_step: int = 1
    skip: List[int] = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class StepRecord:
    step: int
    a: int
    b: int
    ell: float
    theta0_frac: float
    iters: int
    substeps: int
    E_mem: float
    E_ben: float
    E_hin: float
    theta_mean_deg: float
    theta_max_deg: float
    jump_over_L: float
    change_over_L: float
    guard_retry: int
    line: int
    neg_modes: Optional[int] = None


class Backend:
    """What the making needs from a plate (device or, in tests, the oracle).

    rebuild(cuts)            shape functions of every domain with these creases
    set_fold_state(f, th, a) per crease GP rest angle and activity of fold f
    face_displacements(x)    (ng, 3) displacement at the face GPs
    energies(x)              (E_mem, E_ben, E_hin)
    fold_angles(f, x)        theta at fold f's crease GPs
    solve(x)                 LineSearch at lambda = 0 from x (in place):
                             (converged, Newton attempts)
    fold_gps(f)              (gx, gy, w) of fold f's left crease GPs
    theta_star(f)            fold_sign * angle_target of fold f
    """


def run(backend, x: np.ndarray, lines: List[Line], static_creases: List[Crease], cfg: CreasingConfig,
        log: Optional[Callable[[str], None]] = None,
        on_step: Optional[Callable[[StepRecord, np.ndarray], None]] = None) -> List[StepRecord]:
    """Progressive::run. x (the start, e.g. the initial folded condition) is
    updated in place to the last accepted state; returns one record per
    accepted step. On entry the shape functions must be those of every cut in
    full. on_step(record, x) runs after each accepted step (the reference's
    on_converged export)."""
    say = log or (lambda s: None)
    if not lines:
        raise ValueError("Progressive creasing: no line to make")
    nlines = len(lines)
    Sl = [cumulative_arc_length(ln.full) for ln in lines]
    Nl = [len(ln.full.vx) - 1 for ln in lines]
    if min(Nl) < 1:
        raise ValueError("Progressive creasing: crease path has no segments")
    Lc = sum(S[-1] for S in Sl)
    gp_s, theta0, active, weights = {}, {}, {}, {}
    all_folds = []
    for l, ln in enumerate(lines):
        for f in ln.folds:
            gx, gy, w = backend.fold_gps(f)
            gp_s[f] = np.array([project_arc_length(px, py, ln.full, Sl[l]) for px, py in zip(gx, gy)])
            theta0[f] = np.zeros(gx.size)
            active[f] = np.zeros(gx.size, np.int8)
            weights[f] = np.asarray(w)
            all_folds.append(f)
    multi = len(all_folds) > 1

    st = []
    for l, ln in enumerate(lines):
        for s in schedule(ln.mode, ln.tip_step, cfg.num_steps, Nl[l], ln.skip):
            empty = (s.a == 0 and s.b == Nl[l]) if s.outside else (s.b == s.a)
            if l > 0 and empty:
                continue
            st.append((l, s))

    records: List[StepRecord] = []
    prev_E_tot = 0.0
    prev_edge_ends, prev_edge_line = 0, 0
    prev = (nlines - 1, 0, Nl[-1], False)
    prev_active = {f: np.zeros_like(active[f]) for f in all_folds}

    def max_distance(u1, u2):
        return float(np.sqrt(((u1 - u2) ** 2).sum(axis=1)).max()) if u1.shape[0] else 0.0

    for k, (l, s) in enumerate(st):
        ln = lines[l]
        full, S, N = ln.full, Sl[l], Nl[l]
        whole = s.outside and s.a >= s.b
        ell = (S[N] if whole else S[s.a] + (S[N] - S[s.b])) if s.outside else S[s.b] - S[s.a]
        ell += sum(Sl[q][-1] for q in range(l))

        # 1. move the tips: rebuild the shape functions, report the jump
        jump = 0.0
        if (l, s.a, s.b, s.outside) != prev:
            u_old = backend.face_displacements(x)
            cuts = list(static_creases) + [lines[q].full for q in range(l)]
            if not s.outside:
                if s.b > s.a:
                    cuts.append(truncate(full, s.a, s.b))
            elif whole:
                cuts.append(full)
            else:
                if s.a > 0:
                    cuts.append(truncate(full, 0, s.a))
                if s.b < N:
                    cuts.append(truncate(full, s.b, N))
            backend.rebuild(cuts)
            jump = max_distance(u_old, backend.face_displacements(x)) / cfg.L_ref
            prev = (l, s.a, s.b, s.outside)

        # 2. creases on: completed lines in full, this line between its tips
        newest = []
        for q, lq in enumerate(lines):
            for f in lq.folds:
                th_star = backend.theta_star(f)
                for i in range(gp_s[f].size):
                    frac = 1.0
                    if q < l:
                        on = True
                    elif q > l:
                        on = False
                    elif not s.outside:
                        on = s.b > s.a and S[s.a] < gp_s[f][i] < S[s.b]
                        frac = s.frac
                    else:
                        on = whole or gp_s[f][i] < S[s.a] or gp_s[f][i] > S[s.b]
                        frac = s.frac
                    active[f][i] = 1 if on else 0
                    theta0[f][i] = frac * th_star if on else 0.0
                    if on and not prev_active[f][i]:
                        newest.append((f, i))
        say(f"--- step {k}/{len(st) - 1} | line {l} | vertices {'outside ' if s.outside else ''}"
            f"[{s.a}, {s.b}] | ell/L_c = {ell / Lc:.4f} | theta0 frac = {s.frac:.4f} | jump = {jump:.3e} L")

        if s.outside:
            edge_ends = (int((s.a > 0 or whole) and full.ext_s) + int((s.b < N or whole) and full.ext_e))
        else:
            edge_ends = (int(s.a == 0 and full.ext_s) + int(s.b == N and full.ext_e)) if s.b > s.a else 0
        topology_step = edge_ends != prev_edge_ends or l != prev_edge_line
        motion_limit = cfg.guard_max_motion_topology if topology_step else cfg.guard_max_motion

        def push_state():
            for f in all_folds:
                backend.set_fold_state(f, theta0[f], active[f])

        def evaluate(xs, u0):
            r = {"change": max_distance(u0, backend.face_displacements(xs)) / cfg.L_ref}
            r["E_mem"], r["E_ben"], r["E_hin"] = backend.energies(xs)
            means, maxs, wts = [], [], []
            for f in all_folds:
                m, mx = active_fold_angles(backend.fold_angles(f, xs), weights[f], active[f])
                means.append(m); maxs.append(mx)
                wts.append(float(np.sum(weights[f][active[f] != 0])))
            if not multi:
                r["th_mean"], r["th_max"] = means[0], maxs[0]
            else:
                W = sum(wts)
                r["th_mean"] = sum(w * abs(m) for w, m in zip(wts, means)) / W if W > 0 else 0.0
                r["th_max"] = max(maxs)
            return r

        def guard_why(r):
            why = ""
            E_tot = r["E_mem"] + r["E_ben"] + r["E_hin"]
            if r["th_max"] > cfg.guard_max_theta_deg:
                why += f" max fold angle {r['th_max']:.2f} deg > {cfg.guard_max_theta_deg:.1f};"
            if prev_E_tot > 1e-10 and E_tot > cfg.guard_energy_factor * prev_E_tot:
                why += f" total energy {prev_E_tot:.3e} -> {E_tot:.3e};"
            if r["change"] > motion_limit:
                why += f" step motion {r['change']:.3f} L > {motion_limit:.3f} L;"
            return why

        def set_newest(fac):
            for f, i in newest:
                theta0[f][i] = fac * s.frac * backend.theta_star(f)
            push_state()

        # 3. solve: Newton at the full target, then the newest-point theta0 ramp
        x0 = x.copy()
        u0 = backend.face_displacements(x)
        set_newest(1.0)
        substeps, retry, why = 0, 0, ""
        ok, iters = backend.solve(x)
        r = None
        if ok:
            r = evaluate(x, u0)
            why = guard_why(r)
            if why:
                say(f"!!! GUARD: step {k} converged to a refused state:{why}")
                retry = 1
        else:
            say(">> Newton failed.")
        if not ok or why:
            ok = False
            if newest and cfg.fallback_substeps > 0:
                x[:] = x0
                ok = True
                for j in range(1, cfg.fallback_substeps + 1):
                    set_newest(j / cfg.fallback_substeps)
                    say(f">> fallback sub-step {j}/{cfg.fallback_substeps}")
                    okj, it = backend.solve(x)
                    iters += it
                    if not okj:
                        ok = False
                        why = ""
                        break
                substeps = cfg.fallback_substeps
                if ok:
                    r = evaluate(x, u0)
                    why = guard_why(r)
                    ok = not why
        if not ok:
            say(f"!!! PROGRESSIVE CREASING STOPPED at step {k}" + (f" by the guard:{why}" if why else
                                                                    ": no equilibrium"))
            x[:] = x0
            break
        for f in all_folds:
            prev_active[f] = active[f].copy()
        prev_edge_ends, prev_edge_line = edge_ends, l
        nm = backend.neg_modes() if hasattr(backend, "neg_modes") else None
        rec = StepRecord(k, s.a, s.b, ell, s.frac, iters, substeps, r["E_mem"], r["E_ben"], r["E_hin"],
                         r["th_mean"], r["th_max"], jump, r["change"], retry, l, nm)
        say(f" >> step {k}: iters {iters} (guard retry {retry}, sub-steps {substeps}) | E_mem {rec.E_mem:.4e} "
            f"E_ben {rec.E_ben:.4e} E_hin {rec.E_hin:.4e} | theta mean {rec.theta_mean_deg:.4f} max "
            f"{rec.theta_max_deg:.4f} deg | step change {rec.change_over_L:.3e} L")
        records.append(rec)
        prev_E_tot = rec.E_mem + rec.E_ben + rec.E_hin
        if on_step is not None:
            on_step(rec, x)
    return records


# ---------------------------------------------------------------------------
# device backend and the driver's set-up
# ---------------------------------------------------------------------------

class DeviceBackend(Backend):
    """The device plate (curvami_gpu.Plate built by plate_from_sheet) with the
    LineSearch of continuation.ContinuationSolver over continuation.PlateProblem."""

    def __init__(self, plate, sheet, solver, exact_hinge: bool = True, theta_star=None):
        self.plate, self.sheet, self.solver = plate, sheet, solver
        self.exact_hinge = exact_hinge
        self._theta_star = list(theta_star) if theta_star is not None else [f.Ang0 for f in sheet.folds]
        self._psi = None
        self._plan_stale = False

    def fold_gps(self, f):
        g = self.sheet.folds[f].left
        return g.gx, g.gy, g.w

    def theta_star(self, f):
        return self._theta_star[f]

    def rebuild(self, cuts):
        p = self.plate
        p.clear_creases()
        for cr in cuts:
            p.add_crease(cr.vx, cr.vy, cr.ext_s, cr.ext_e, cr.closed)
        p.shape_functions_all(self.sheet.order, self.sheet.kernel)
        self._psi = None

    def set_fold_state(self, f, theta0, active):
        self.plate.set_fold_gp_state(f, theta0, active)
        self._plan_stale = True

    def face_displacements(self, x):
        if self._psi is None:
            off, ids, _ = self.plate.get_neighbours(self.plate.face_dom)
            psi = self.plate.get_shape_functions(self.plate.face_dom)[0]
            self._psi = (np.repeat(np.arange(off.size - 1), np.diff(off)), ids, psi, off.size - 1)
        g, ids, psi, ng = self._psi
        X = np.asarray(x).reshape(-1, 3)
        return np.stack([np.bincount(g, psi * X[ids, c], ng) for c in range(3)], axis=1)

    def energies(self, x):
        return self.plate.energies(x)

    def fold_angles(self, f, x):
        return self.plate.fold_angles(f, x)

    def solve(self, x):
        from continuation import PlateProblem
        if self._plan_stale:
            self.plate.init_assembly()
            self._plan_stale = False
        prob = PlateProblem(self.plate, flag_rig=False, exact_hinge=self.exact_hinge)
        rep = self.solver.line_search(prob, 0.0, x)
        return rep.converged, rep.attempts

    def neg_modes(self):
        """Negative modes of the reduced Hessian from the last KKT factorization
        (negative pivots minus the independent constraint rows,
        ProgressiveCreasing.h:151-153), or None without an inertia."""
        neg = getattr(self.solver.kkt, "neg_pivots", -1)
        return None if neg < 0 else int(neg) - int(self.solver.ind_rows.size)


def lines_from_config(ms, pc: dict) -> (List[Line], List[Crease]):
    """The lines to make and the static cuts (main.cpp:558-606): each line is
    one cut carrying one or more folds; a line's cut is removed from the static
    cuts, which must hold it exactly."""
    from mesh_setup import crease_from_mesh
    sh = ms.sheet
    tags_of_fold = [fi.crease_tag for fi in ms.fold_info]
    line_cfgs = pc["lines"] if "lines" in pc else [{"crease_tags": [int(pc["crease_tag"])],
                                                    "mode": pc.get("mode", "one_end"),
                                                    "tip_step": pc.get("tip_step", 1)}]
    static = list(sh.creases)
    lines = []
    for lc in line_cfgs:
        tags = [int(t) for t in lc["crease_tags"]]
        folds = []
        for t in tags:
            if t not in tags_of_fold:
                raise ValueError(f"creasing: crease tag {t} is not one of the folds")
            fi = len(tags_of_fold) - 1 - tags_of_fold[::-1].index(t)
            if sh.folds[fi].rigid:
                raise ValueError("a progressive crease must be elastic (is_rigid false)")
            folds.append(fi)
        full = crease_from_mesh(ms.input.mesh, tags)
        hit = [i for i, c in enumerate(static) if c.vx.size == full.vx.size and np.array_equal(c.vx, full.vx)
               and np.array_equal(c.vy, full.vy)]
        if not hit:
            raise ValueError("each progressive line must be one cut: list exactly its crease tags as one "
                             "\"cuts\" entry (or omit \"cuts\" for a single crease)")
        static.pop(hit[0])
        mode = lc.get("mode", "one_end")
        if "start" in lc:
            px, py = lc["start"]
            d0 = math.hypot(full.vx[0] - px, full.vy[0] - py)
            d1 = math.hypot(full.vx[-1] - px, full.vy[-1] - py)
            mode = "one_end" if d0 <= d1 else "one_end_reverse"
        lines.append(Line(folds, full, mode, int(lc.get("tip_step", 1)), list(lc.get("skip", []))))
    return lines, static
