"""The bench workload at full size on the device: the 201 x 201 arc-crease sheet
(640,000 face GPs, 28.2M pairs, rho = 3.75 h, quadratic basis, elastic hinge),
with the quartic window and with the cubic B-spline window, checked

- against the oracle on a bounded sample of the same sheet: the face GPs of
  cell rows 95-104 (32,000 GPs, crossed by the arc) have the neighbour lists
  (bit-exact) and shape functions (the parity tests' per-GP bars) that the
  oracle computes for them;
- through size-independent properties on every GP: the reproducing conditions
  of the quadratic basis, a residual and energies that vanish under a rigid
  motion, F_int = dE/dx, a tangent consistent with central differences of the
  residual, bit-identical repeats, and the FP32 tangent within its stated bar
  of the FP64 one.

Bars are stated per check. Reproducing conditions: per GP
max(1e-11, 1e3 eps cond(M)) for the values and, scaled by rho and rho^2,
max(1e-10, 1e4 eps cond(M)) for the derivatives (sums of ~44 products of the
shape functions, each within its own cond-scaled bar).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import curvami_gpu as G
import oracle as O
import workloads as W
from _cases import EPS, moment_cond, normwise, pair_gp, sf_tolerance, sheet

pytestmark = pytest.mark.gpu

Q = 16                     # GPs per cell (4 x 4)
ROWS = (95, 105)           # cell rows of the oracle sample


@pytest.fixture(scope="module", params=["arc201", "arc201_bspline"])
def full(request):
    """The sheet with the quartic window (the reference's) and with the cubic
    B-spline window (the configurations' window)."""
    sh = sheet(request.param)
    print(f"{sh.name}: {sh.nn} nodes, {sh.face.n} face GPs", flush=True)
    return sh, G.plate_from_sheet(sh)


@pytest.fixture(scope="module")
def face(full):
    sh, gpu = full
    off, ids, s = gpu.get_neighbours(0)
    sf = gpu.get_shape_functions(0)
    cond = moment_cond(sh.face.gx, sh.face.gy, sh.X, sh.Y, sh.rho, off, ids, sf[0], sh.order, sh.kernel)
    print(f"pairs {off[-1]}, cond(M) median {np.median(cond):.3e}, max {cond.max():.3e}", flush=True)
    return off, ids, s, sf, cond


def test_sample_matches_oracle(full, face):
    sh, _ = full
    off, ids, s, sf, _ = face
    ncell = sh.n_side - 1
    lo, hi = ROWS[0] * ncell * Q, ROWS[1] * ncell * Q
    sub = W.GPSet(sh.face.gx[lo:hi], sh.face.gy[lo:hi], sh.face.w[lo:hi])
    sample = W.Sheet(sh.name + "_rows95-104", sh.n_side, sh.X, sh.Y, sub, sh.creases, sh.folds, sh.rho,
                     sh.order, sh.kernel, sh.E, sh.nu, sh.t)
    orc = O.plate_from_sheet(sample, compute=False)
    o_off, o_ids, o_s = orc.neighbours(0, sample.rho)
    p0, p1 = off[lo], off[hi]
    assert np.array_equal(off[lo:hi + 1] - p0, o_off)
    assert np.array_equal(ids[p0:p1], o_ids)
    assert np.array_equal(s[p0:p1].view(np.int64), o_s.view(np.int64))
    st, _, _ = orc.shape_functions(0, sample.rho, sample.order, sample.kernel, 16)
    assert st == 0
    ref = orc.get_shape_functions(0, o_off[-1])
    got = [q[p0:p1] for q in sf]
    g = pair_gp(o_off)
    ng = o_off.size - 1
    cond = moment_cond(sub.gx, sub.gy, sh.X, sh.Y, sh.rho, o_off, o_ids, ref[0], sh.order, sh.kernel)
    errs = (normwise([got[0]], [ref[0]], g, ng), normwise(got[1:3], ref[1:3], g, ng),
            normwise(got[3:6], ref[3:6], g, ng))
    for e, tol, nm in zip(errs, sf_tolerance(cond), ("psi", "grad", "hess")):
        print(f"sample {nm}: max {e.max():.3e}, median {np.median(e):.3e}")
        assert np.all(e <= tol), f"{nm}: {int(np.sum(e > tol))} GPs above their bar"


def test_reproducing_conditions(full, face):
    """sum_I D Psi_I(x) xi_I^a eta_I^b (xi = (X_I - x) / rho) equals the
    derivative D of (x' - x)^a (y' - y)^b at x' = x, scaled: 1 for Psi and
    (a, b) = (0, 0), rho^-1 for d/dx and (1, 0), 2 rho^-2 for d2/dx2 and
    (2, 0), rho^-2 for d2/dxdy and (1, 1), 0 otherwise."""
    sh, _ = full
    off, ids, _, sf, cond = face
    g = pair_gp(off)
    ng = off.size - 1
    xi = (sh.X[ids] - sh.face.gx[g]) / sh.rho
    eta = (sh.Y[ids] - sh.face.gy[g]) / sh.rho
    mons = {(a, b): xi ** a * eta ** b for a in range(3) for b in range(3 - a)}
    # (plane, scale, the monomial whose sum is nonzero, its value after scaling)
    planes = [(0, 1.0, (0, 0), 1.0), (1, sh.rho, (1, 0), 1.0), (2, sh.rho, (0, 1), 1.0),
              (3, sh.rho ** 2, (2, 0), 2.0), (4, sh.rho ** 2, (0, 2), 2.0), (5, sh.rho ** 2, (1, 1), 1.0)]
    for q, scale, hit, val in planes:
        worst = np.zeros(ng)
        for ab, m in mons.items():
            r = np.bincount(g, sf[q] * scale * m, ng) - (val if ab == hit else 0.0)
            worst = np.maximum(worst, np.abs(r))
        bar = np.maximum(1e-11, 1e3 * EPS * cond) if q == 0 else np.maximum(1e-10, 1e4 * EPS * cond)
        print(f"plane {q}: max error {worst.max():.3e} (median {np.median(worst):.3e})")
        assert np.all(worst <= bar), f"plane {q}: {int(np.sum(worst > bar))} GPs above their bar"


def _rigid(sh, angle=0.3, shift=(0.2, -0.1, 0.05)):
    x = np.zeros(3 * sh.nn)
    c, s = np.cos(angle), np.sin(angle)
    x[0::3] = sh.X * c - sh.Y * s - sh.X + shift[0]
    x[1::3] = sh.X * s + sh.Y * c - sh.Y + shift[1]
    x[2::3] = shift[2]
    return x


def test_rigid_motion_is_stress_free(full):
    sh, gpu = full
    x = W.displacement(sh.nn, 5, 1e-3)
    F_gen, _ = gpu.assemble(x, tangent=False)
    E_gen = sum(gpu.energies(x))
    F_rig, _ = gpu.assemble(_rigid(sh), tangent=False)
    E_rig = gpu.energies(_rigid(sh))
    print(f"rigid motion: |F| {np.abs(F_rig).max():.3e} (generic {np.abs(F_gen).max():.3e}), energies {E_rig}")
    assert np.abs(F_rig).max() <= 1e-9 * np.abs(F_gen).max()
    assert max(abs(e) for e in E_rig) <= 1e-9 * E_gen


def test_residual_is_energy_gradient(full):
    sh, gpu = full
    x = W.displacement(sh.nn, 7, 2e-3)
    F, _ = gpu.assemble(x, tangent=False)
    h = 1e-7
    for seed in (1, 2):
        v = W.displacement(sh.nn, 100 + seed, 1.0)
        fd = (sum(gpu.energies(x + h * v)) - sum(gpu.energies(x - h * v))) / (2 * h)
        print(f"dE/dx.v {fd:.9e}, F.v {F @ v:.9e}")
        assert abs(fd - F @ v) <= 1e-8 * np.linalg.norm(F) * np.linalg.norm(v)


def _upper_view(gpu, K, with_folds):
    outer, inner = gpu.pattern(with_folds=with_folds)
    n = outer.size - 1
    U = sp.csc_matrix((K, inner, outer), shape=(n, n))
    return sp.triu(U) + sp.triu(U, 1).T       # selfadjointView<Upper>


@pytest.mark.parametrize("elastic", [False, True])
def test_tangent_consistent_with_residual(full, elastic):
    """Face terms only, and face + hinge with the exact (FD-Hessian) hinge
    tangent: K v against central differences of F_int along random unit
    directions, consistent to < 1e-6 (the reference's TangentCheck verdict)."""
    sh, gpu = full
    x = W.displacement(sh.nn, 3, 1e-3)
    F, K = gpu.assemble(x, elastic_folds=elastic, exact_hinge=elastic)
    Kf = _upper_view(gpu, K, elastic)
    rng = np.random.default_rng(9)
    h = 1e-6 * max(1.0, np.abs(x).max())
    for _ in range(2):
        v = rng.standard_normal(x.size); v /= np.linalg.norm(v)
        Fp, _ = gpu.assemble(x + h * v, elastic_folds=elastic, tangent=False)
        Fm, _ = gpu.assemble(x - h * v, elastic_folds=elastic, tangent=False)
        fd = (Fp - Fm) / (2 * h)
        rel = np.linalg.norm(Kf @ v - fd) / np.linalg.norm(fd)
        print(f"elastic={elastic}: |K v - FD| / |FD| = {rel:.3e}")
        assert rel < 1e-6


def test_repeats_bit_identical(full, face):
    sh, gpu = full
    sf0 = face[3]
    gpu.shape_functions_all(sh.order, sh.kernel)
    sf1 = gpu.get_shape_functions(0)
    for a, b in zip(sf0, sf1):
        assert np.array_equal(a.view(np.int64), b.view(np.int64))
    x = W.displacement(sh.nn, 11, 1e-3)
    F1, K1 = gpu.assemble(x)
    F2, K2 = gpu.assemble(x)
    assert np.array_equal(F1.view(np.int64), F2.view(np.int64))
    assert np.array_equal(K1.view(np.int64), K2.view(np.int64))


def test_fp32_tangent_within_bar(full):
    """The mixed path's tangent against the FP64 tangent: <= 1e-5 of max|K|
    (max and Frobenius), the residual unchanged to FMA round-off."""
    sh, gpu = full
    x = W.displacement(sh.nn, 1, 1e-3)
    F64, K64 = gpu.assemble(x)
    F32, K32 = gpu.assemble(x, fp32_tangent=True)
    assert np.abs(F64 - F32).max() <= 1e-13 * np.abs(F64).max()
    ek = np.abs(K32 - K64).max() / np.abs(K64).max()
    ef = np.linalg.norm(K32 - K64) / np.linalg.norm(K64)
    print(f"FP32 tangent: max rel {ek:.3e}, Frobenius rel {ef:.3e}")
    assert ek <= 1e-5 and ef <= 1e-5
