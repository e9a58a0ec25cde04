"""Writes progressive_smoke.npz: the oracle's progressive-creasing run of the
reference's smoke example (smoke.json on CB_n20.msh: cubic basis, rho = 6 x
median 3rd-NN spacing, crease grown middle-out in 5 steps to theta0* = 0.5 deg,
LineSearch with tol_force 1e-10, tol_energy 1e-12, exact hinge tangent),
through the oracle backend of tests/_progressive_oracle.py.

Stored per accepted step: tip vertices a, b, Newton attempts, E_mem, E_ben,
E_hin, mean and largest crease angle (deg); and the final displacement.
Run from the repository root:
    python tests/golden/make_progressive_fixture.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "curvami-official_amd"), os.path.join(ROOT, "tests")]

from _progressive_oracle import run_oracle_smoke  # noqa: E402


def main():
    t0 = time.time()
    recs, x = run_oracle_smoke(os.path.join(HERE, "reference_inputs"), log=print)
    print(f"{time.time() - t0:.1f} s")
    cols = ("a", "b", "iters", "E_mem", "E_ben", "E_hin", "theta_mean_deg", "theta_max_deg")
    np.savez_compressed(os.path.join(HERE, "progressive_smoke.npz"), x=x,
                        **{c: np.array([getattr(r, c) for r in recs]) for c in cols})


if __name__ == "__main__":
    main()
