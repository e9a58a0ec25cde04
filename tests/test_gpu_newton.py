"""Constraint rows, external load and the Newton load path on the device,
against the oracle.

Bars: the triplet pattern bit-identical; constraint values, Jacobian values and
the external load within the shape-function round-off (they are linear in Psi,
which the device and the oracle compute to ~1e-13, not bit-identically); the converged
displacement of the 20-increment load path within 1e-8 (inf-norm, relative)
of the oracle's, with every increment converged (golden fixture written by
tests/golden/make_newton_fixture.py); four arc-length steps with the device
assembly against the oracle's (the same steps, the state within 1e-8). The
FP32-tangent (mixed) path, finished by one FP64-tangent step per increment,
reports its Newton iteration counts and converges to the FP64 oracle's state
within 1e-8.
"""
import os
import time

import numpy as np
import pytest

import continuation as CS
import curvami_gpu as G
import oracle as O
import workloads as W
from _cases import RHO_FACTOR

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "newton_straight21_moment.npz")


@pytest.fixture(scope="module")
def c1m():
    sh = W.with_edge_moment_load(W.config1(rho_factor=RHO_FACTOR))
    return sh, O.plate_from_sheet(sh, nthreads=4), G.plate_from_sheet(sh)


@pytest.mark.parametrize("rigid", [False, True])
def test_constraint_rows_match_oracle(c1m, rigid):
    sh, orc, gpu = c1m
    x = W.rigid_fold_displacement(sh, 0.5, seed=3, sigma=1e-3)
    rows, cols = gpu.constraint_layout(rigid=rigid)
    c_o, r_o, col_o, v_o = orc.constraints(x, flag_rig=rigid)
    assert np.array_equal(rows, r_o) and np.array_equal(cols, col_o), "triplet pattern differs"
    c_g, v_g = gpu.constraints(x)
    assert c_g.shape == c_o.shape and v_g.shape == v_o.shape
    ec = np.abs(c_g - c_o).max()
    ev = np.abs(v_g - v_o).max()
    print(f"rigid={rigid}: {c_o.size} rows, {v_o.size} triplets, max|dc| {ec:.2e} (max|c| {np.abs(c_o).max():.2e}), "
          f"max|dJ| {ev:.2e} (max|J| {np.abs(v_o).max():.2e})")
    # c is a Psi-weighted sum of |x| <= ~3e-3 (linear rows) or an angle (rigid rows)
    assert ec <= 1e-12
    assert ev <= 1e-10 * max(1.0, np.abs(v_o).max())


def test_external_force(c1m):
    sh, orc, gpu = c1m
    Fo = orc.external_force()
    Fg = gpu.external_force()
    assert np.abs(Fo).max() > 0
    e = np.abs(Fg - Fo).max() / np.abs(Fo).max()
    print(f"F_ext: max rel {e:.2e}")
    assert e <= 1e-11


def prob_kkt_stats(prob):
    return getattr(prob, "kkt_stats", {})


def _run_path(gpu, fp32):
    ref = np.load(FIXTURE)
    prob = CS.PlateProblem(gpu, flag_rig=False, fp32_tangent=fp32)
    solver = CS.ContinuationSolver()
    solver.tol_force = float(ref["tol_force"])
    solver.tol_energy = float(ref["tol_energy"])
    t0 = time.perf_counter()
    res = CS.load_path(prob, increments=W.NEWTON_INCREMENTS, lam_total=1.0, solver=solver)
    dt = time.perf_counter() - t0
    prob.kkt_stats = solver.kkt.stats()
    return ref, res, prob, dt


def test_newton_load_path_matches_oracle(c1m):
    sh, _, gpu = c1m
    ref, res, prob, dt = _run_path(gpu, fp32=False)
    err = np.abs(res.x - ref["x"]).max() / np.abs(ref["x"]).max()
    print(f"FP64 path: {len(res.converged)} increments, iterations {res.iterations} "
          f"(oracle {ref['iterations'].tolist()}), |dx|/|x| {err:.3e}, {dt:.1f} s, "
          f"{prob.assemblies} tangents, {prob.residuals} residuals")
    assert res.all_converged and len(res.converged) == W.NEWTON_INCREMENTS == len(ref["iterations"])
    assert list(res.iterations) == ref["iterations"].tolist()
    assert err <= 1e-8
    st = prob_kkt_stats(prob)
    print(f"KKT: {st}")


def test_newton_mixed_precision_tangent(c1m):
    """FP32 face tangent (inexact Newton), FP64 residual and constraint rows. The
    FP32 iteration converges linearly and stops just inside the stopping rule
    (max|P R| <= tol_force), where the quadratically converging FP64 iteration
    overshoots it by orders of magnitude; one Newton step on the FP64 tangent
    once the rule holds (ContinuationSolver.finish_fp64) brings the state to
    the FP64 one, and a stalled FP32 step switches the rest of the increment to
    the FP64 tangent. Bar: 1e-8 relative to the FP64 oracle's path; the
    iteration counts (FP32 and FP64 tangents) and their change are reported."""
    sh, _, gpu = c1m
    ref, res, prob, dt = _run_path(gpu, fp32=True)
    err = np.abs(res.x - ref["x"]).max() / np.abs(ref["x"]).max()
    d_it = int(np.sum(res.iterations) - np.sum(ref["iterations"]))
    print(f"mixed path: iterations {res.iterations} (sum {sum(res.iterations)}: {sum(res.iterations) - sum(res.fp64_steps)} "
          f"on the FP32 tangent, {sum(res.fp64_steps)} on the FP64 one; oracle FP64 sum "
          f"{int(np.sum(ref['iterations']))}, delta {d_it:+d}), |dx|/|x| {err:.3e}, {dt:.1f} s")
    assert res.all_converged and len(res.converged) == W.NEWTON_INCREMENTS
    assert all(f >= 1 for f in res.fp64_steps)      # every increment ends on the FP64 tangent
    assert err <= 1e-8


def test_arc_length_matches_oracle(c1m):
    """ArcLength (ContinuationSolver.h:329-538) on the moment-loaded straight-crease
    sheet, four steps from x = 0 with psi = 1, ds = 0.05, ds_max = 0.5: the path
    with the device assembly and the oracle's (oracle/newton.py) take the same
    steps (lambda within 1e-8 relative, equal corrector iteration counts) to the
    same state (1e-8, inf-norm, relative)."""
    import newton as ON
    sh, orc, gpu = c1m
    n = 3 * sh.nn
    x_o, lam_o, st_o = ON.arc_length(ON.OracleProblem(orc, flag_rig=False, nthreads=8), 0.0, np.zeros(n), 1.0,
                                     ds=0.05, dsmax=0.5, kmax=4)
    s = CS.ContinuationSolver()
    s.psi, s.ds, s.dsmax, s.kmax = 1.0, 0.05, 0.5, 4
    rep = s.arc_length(CS.PlateProblem(gpu, flag_rig=False), 0.0, np.zeros(n), 1.0)
    print(f"arc length: device {[(round(a.lam, 6), a.iterations) for a in rep.steps]}, "
          f"oracle {[(round(float(l), 6), it) for _, l, it in st_o]}")
    assert len(rep.steps) == len(st_o) == 4
    for a, (_, lam, it) in zip(rep.steps, st_o):
        assert a.iterations == it and abs(a.lam - lam) <= 1e-8 * abs(lam)
    assert np.abs(rep.x - x_o).max() <= 1e-8 * np.abs(x_o).max()


def test_ls_miura_line_search_matches_oracle():
    """The reference's regression configuration ls_miura (four rigid folds
    pre-set to +-15.26 / 30 deg, two cuts, corner supports; cubic basis,
    rho = 6h): the initial folded condition and one LineSearch at target_load 0
    with the rigid-fold rows, on the device, against the oracle's run
    (tests/golden/lsmiura_linesearch.npz, make_lsmiura_fixture.py). Bars: x0
    and the converged x within 1e-8 (inf-norm, relative), the same number of
    Newton iterations."""
    import json
    import mesh_setup as MS
    inputs = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_inputs")
    ref = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lsmiura_linesearch.npz"))
    with open(os.path.join(inputs, "ls_miura.json")) as f:
        cfg = json.load(f)
    ms = MS.sheet_from_config(cfg, base_dir=inputs, mesh_file=os.path.join(inputs, "Miura.msh"))
    plate = G.plate_from_sheet(ms.sheet)
    prob = CS.PlateProblem(plate, flag_rig=True)
    solver = CS.ContinuationSolver()
    t0 = time.perf_counter()
    x0 = CS.initial_condition_with_folds(prob, solver)
    e0 = np.abs(x0 - ref["x0"]).max() / np.abs(ref["x0"]).max()
    x = x0.copy()
    rep = solver.line_search(prob, float(cfg.get("target_load", 0.0)), x)
    dt = time.perf_counter() - t0
    e = np.abs(x - ref["x"]).max() / np.abs(ref["x"]).max()
    print(f"ls_miura: x0 rel {e0:.2e}; converged {rep.converged} in {rep.iterations} iterations "
          f"(oracle {int(ref['iterations'])}), x rel {e:.2e}, {dt:.1f} s; energies {plate.energies(x)} "
          f"(oracle {ref['energies']})")
    assert e0 <= 1e-8 and rep.converged and rep.iterations == int(ref["iterations"]) and e <= 1e-8
