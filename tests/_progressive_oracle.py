"""The progressive-creasing backend over the oracle plate (test
infrastructure): the same stepping as on the device, with the CPU
restatement's shape functions, residual, tangent, constraint rows, energies
and crease angles, and the oracle's LineSearch (oracle/newton.py)."""
import numpy as np

import newton as ON
import oracle as O
from progressive import Backend


class OracleBackend(Backend):
    def __init__(self, orc, sheet, theta_star, exact_hinge=True, tol_force=1e-8, tol_energy=1e-8,
                 nthreads=8):
        self.orc, self.sheet = orc, sheet
        self._theta_star = list(theta_star)
        self.exact_hinge = exact_hinge
        self.tol = dict(tol_force=tol_force, tol_energy=tol_energy)
        self.nthreads = nthreads
        off, ids, _ = orc.neighbours(orc.face_dom, sheet.rho)
        self._g = np.repeat(np.arange(off.size - 1), np.diff(off))
        self._ids, self._ng, self._np = ids, off.size - 1, int(off[-1])

    def fold_gps(self, f):
        g = self.sheet.folds[f].left
        return g.gx, g.gy, g.w

    def theta_star(self, f):
        return self._theta_star[f]

    def rebuild(self, cuts):
        o = self.orc
        o.clear_creases()
        for cr in cuts:
            o.add_crease(cr.vx, cr.vy, cr.ext_s, cr.ext_e, cr.closed)
        for d in range(len(o.domains)):
            st, nb, fb = o.shape_functions(d, self.sheet.rho, self.sheet.order, self.sheet.kernel, self.nthreads)
            if st != 0:
                raise RuntimeError(f"oracle RKPM status {st} on domain {d}")

    def set_fold_state(self, f, theta0, active):
        self.orc.set_fold_gp_state(f, theta0, active)

    def face_displacements(self, x):
        psi = self.orc.get_shape_functions(self.orc.face_dom, self._np)[0]
        X = np.asarray(x).reshape(-1, 3)
        return np.stack([np.bincount(self._g, psi * X[self._ids, c], self._ng) for c in range(3)], axis=1)

    def energies(self, x):
        return self.orc.energies(x)

    def fold_angles(self, f, x):
        return self.orc.fold_angles(f, x)

    def solve(self, x):
        prob = ON.OracleProblem(self.orc, flag_rig=False, nthreads=self.nthreads, exact_tangent=self.exact_hinge)
        ok, it = ON.line_search(prob, 0.0, x, **self.tol)
        return ok, it


def smoke_case(inputs_dir):
    """The reference's smoke example (examples/smoke_centre_crease/smoke.json on
    CB_n20.msh): configuration and the sheet built from it."""
    import json
    import os

    import mesh_setup as MS
    with open(os.path.join(inputs_dir, "smoke.json")) as f:
        cfg = json.load(f)
    ms = MS.sheet_from_config(cfg, base_dir=inputs_dir, mesh_file=os.path.join(inputs_dir, "CB_n20.msh"))
    return cfg, ms


# examples/smoke_centre_crease/README.md:14-20: the last row of the run's step
# log "should read, to round-off" (step, E_mem, E_ben, E_hin, mean theta in deg),
# and every step converges within 5 Newton iterations
SMOKE_KNOWN = {"step": 5, "E_mem": 4.006e-08, "E_ben": 9.404e-06, "E_hin": 1.594e-05, "theta_mean_deg": 0.1962}
SMOKE_MAX_ITERS = 5


def run_oracle_smoke(inputs_dir, log=None):
    import progressive as PR
    cfg, ms = smoke_case(inputs_dir)
    sh = ms.sheet
    orc = O.plate_from_sheet(sh, nthreads=8)
    be = OracleBackend(orc, sh, [fi.sign * fi.angle_target for fi in ms.fold_info],
                       exact_hinge=cfg["solver"].get("fold_exact_tangent", False),
                       tol_force=cfg["solver"]["tol_force"], tol_energy=cfg["solver"]["tol_energy"])
    lines, static = PR.lines_from_config(ms, cfg["creasing"])
    x = np.zeros(3 * sh.nn)
    recs = PR.run(be, x, lines, static, PR.CreasingConfig.from_json(cfg["creasing"]), log=log)
    return recs, x
