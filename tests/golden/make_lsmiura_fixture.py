"""Writes lsmiura_linesearch.npz: the oracle's run of the reference's
regression configuration ls_miura (tests/regression/configs/ls_miura.json on
Miura.msh: four rigid folds pre-set to +-15.26 and 30 deg, two cuts, corner
supports; cubic basis, rho = 6 x median 3rd-NN spacing): the initial folded
condition (NonLinearSolver::GetInitialConditionWithFolds) and one LineSearch
at target_load 0 with the rigid-fold rows (tol_force = tol_energy = 1e-8, the
reference's defaults), both through the oracle plate (oracle.py) and the
oracle's LineSearch (oracle/newton.py).

Stored: x0, the converged x, the Newton iterations, the energies at x.
Run from the repository root:
    python tests/golden/make_lsmiura_fixture.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "curvami-official_amd")]

import continuation as CS  # noqa: E402
import mesh_setup as MS  # noqa: E402
import newton as ON  # noqa: E402
import oracle as O  # noqa: E402

INPUTS = os.path.join(HERE, "reference_inputs")


def main():
    with open(os.path.join(INPUTS, "ls_miura.json")) as f:
        cfg = json.load(f)
    ms = MS.sheet_from_config(cfg, base_dir=INPUTS, mesh_file=os.path.join(INPUTS, "Miura.msh"))
    orc = O.plate_from_sheet(ms.sheet, nthreads=os.cpu_count() or 1)
    prob = ON.OracleProblem(orc, flag_rig=True, nthreads=os.cpu_count() or 1)
    t0 = time.time()
    x0 = CS.initial_condition_with_folds(prob, CS.ContinuationSolver())
    x = x0.copy()
    ok, it = ON.line_search(prob, float(cfg.get("target_load", 0.0)), x, tol_force=1e-8, tol_energy=1e-8)
    print(f"{time.time() - t0:.1f} s, converged {ok} in {it} iterations")
    np.savez_compressed(os.path.join(HERE, "lsmiura_linesearch.npz"), x0=x0, x=x, iterations=it,
                        converged=ok, energies=np.array(orc.energies(x)))


if __name__ == "__main__":
    main()
