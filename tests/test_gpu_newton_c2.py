"""The Newton load path at full size: the 201 x 201 arc-crease sheet (121,203
DOFs, 774 constraint rows: the three-corner supports and the tie rows of the
crease GPs) under the edge-moment load, 20 proportional increments, one
LineSearch each, residual and tangent from the device, the KKT systems by the
host KKT solver (symbolic analysis once, one numerical factorization per
Newton iteration), against the oracle's own path on the same sheet (golden
fixture tests/golden/newton_arc201_moment.npz, written by
make_c2_newton_fixture.py).

Bars: every increment converged, the same Newton iteration count per increment
as the oracle's path, the displacement within 1e-8 (inf-norm, relative) after
increments 1, 10 and 20. The mixed path (FP32 face tangent; a stalled FP32
step switches the increment to the FP64 tangent, and every increment ends with
an FP64 step) converges to the same states within 1e-8 and reports its
iteration counts.

The first two increments run by default; the whole path (about 100
factorizations of the KKT system, minutes on the host) runs with
CVM_LONG_TESTS=1.
"""
import os
import time

import numpy as np
import pytest

import continuation as CS
import curvami_gpu as G
import workloads as W

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "newton_arc201_moment.npz")
LONG = os.environ.get("CVM_LONG_TESTS") == "1"
RHO_FACTOR = 3.75


@pytest.fixture(scope="module")
def c2():
    sh = W.with_edge_moment_load(W.config2(rho_factor=RHO_FACTOR))
    print(f"{sh.name}: {sh.nn} nodes, {sh.face.n} face GPs", flush=True)
    return sh, G.plate_from_sheet(sh), np.load(FIXTURE)


def _path(gpu, ref, fp32, increments):
    """Increments 1..increments of the 20-increment load path; returns the
    iteration counts, the finishing steps, the states after increments 1, 10
    and the last, and the KKT statistics."""
    prob = CS.PlateProblem(gpu, flag_rig=False, fp32_tangent=fp32)
    solver = CS.ContinuationSolver()
    solver.tol_force = float(ref["tol_force"])
    solver.tol_energy = float(ref["tol_energy"])
    x = np.zeros(prob.num_dofs)
    its, f64, states = [], [], {}
    t0 = time.perf_counter()
    for k in range(1, increments + 1):
        rep = solver.line_search(prob, k / W.NEWTON_INCREMENTS, x)
        assert rep.converged, f"increment {k} did not converge"
        its.append(rep.iterations)
        f64.append(rep.fp64_steps)
        if k in (1, 10):
            states[k] = x.copy()
        print(f"increment {k}: {rep.iterations} iterations ({rep.fp64_steps} on the FP64 tangent), "
              f"{time.perf_counter() - t0:.1f} s", flush=True)
    states["last"] = x.copy()
    return its, f64, states, solver.kkt.stats(), time.perf_counter() - t0


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_c2_first_increments_match_oracle(c2):
    sh, gpu, ref = c2
    its, _, states, st, dt = _path(gpu, ref, fp32=False, increments=2)
    e1 = _rel(states[1], ref["x_first"])
    print(f"C2 increments 1-2: iterations {its} (oracle {ref['iterations'][:2].tolist()}), x1 rel {e1:.3e}, "
          f"{dt:.1f} s, KKT {st}")
    assert its == ref["iterations"][:2].tolist()
    assert e1 <= 1e-8


@pytest.mark.skipif(not LONG, reason="the whole 20-increment path takes minutes; run with CVM_LONG_TESTS=1")
def test_c2_load_path_matches_oracle(c2):
    sh, gpu, ref = c2
    assert int(ref["last"]) == W.NEWTON_INCREMENTS, "fixture holds a partial path"
    its, _, states, st, dt = _path(gpu, ref, fp32=False, increments=W.NEWTON_INCREMENTS)
    errs = (_rel(states[1], ref["x_first"]), _rel(states[10], ref["x_10"]), _rel(states["last"], ref["x_last"]))
    print(f"C2 FP64 path: iterations {its} (oracle {ref['iterations'].tolist()}), x rel after 1/10/20 "
          f"{errs[0]:.3e} / {errs[1]:.3e} / {errs[2]:.3e}, {dt:.1f} s, KKT {st}")
    assert its == ref["iterations"].tolist()
    assert max(errs) <= 1e-8


@pytest.mark.skipif(not LONG, reason="the whole 20-increment path takes minutes; run with CVM_LONG_TESTS=1")
def test_c2_mixed_precision_path(c2):
    sh, gpu, ref = c2
    its, f64, states, st, dt = _path(gpu, ref, fp32=True, increments=W.NEWTON_INCREMENTS)
    err = _rel(states["last"], ref["x_last"])
    d_it = int(sum(its) - int(np.sum(ref["iterations"])))
    print(f"C2 mixed path: iterations {its} (sum {sum(its)}: {sum(its) - sum(f64)} on the FP32 tangent, {sum(f64)} on "
          f"the FP64 one; oracle FP64 sum {int(np.sum(ref['iterations']))}, delta {d_it:+d}), x rel after 20 "
          f"{err:.3e}, {dt:.1f} s, KKT {st}")
    assert err <= 1e-8
