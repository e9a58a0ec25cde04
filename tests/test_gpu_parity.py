"""Parity of the HIP path with the CPU oracle on identical seeded inputs.

Bars: neighbour lists and CSC patterns bit-exact; shape functions per GP
normwise (see SF_TOL); F_int to 1e-10 relative (inf-norm); K to 1e-12 of max|K|.
"""
import numpy as np
import pytest

import curvami_gpu as G
import oracle as O
import workloads as W
from _cases import moment_cond, normwise, pair_gp, sf_tolerance, sheet

pytestmark = pytest.mark.gpu

# Shape functions: 1e-12 normwise per GP, relaxed where the moment matrix is
# ill-conditioned (GPs next to a crease). There the reference algorithm itself
# moves by ~1e-11 between builds with and without FMA contraction, so the bar is
# max(1e-12, C * eps * cond(M)) with C = 4 / 16 / 64 for Psi / grad / Hessian.
SF_TOL = 1e-12
# Fraction of face GPs within 1e-12 itself, per derivative order: measured
# 1.0 / 0.9989-1.0 / 0.9836-1.0 on the cases below (the rest sit next to the
# crease, where cond(M) reaches ~1e4).
FRAC_MIN = {"psi": 0.999, "grad": 0.995, "hess": 0.98}


def _build(name):
    sh = sheet(name)
    orc = O.plate_from_sheet(sh)
    gpu = G.plate_from_sheet(sh)
    return sh, orc, gpu


@pytest.fixture(scope="module")
def c1():
    return _build("straight21")


@pytest.fixture(scope="module")
def arc():
    return _build("arc41")


@pytest.fixture(scope="module")
def c1_bs():
    """The straight-crease sheet with the cubic B-spline window."""
    return _build("straight21_bspline")


@pytest.fixture(scope="module")
def arc_bs():
    """The arc-crease sheet with the cubic B-spline window."""
    return _build("arc41_bspline")


@pytest.fixture(scope="module")
def wide():
    """Support 6h (the reference's default support radius, 3rd-NN distance x
    (order + 3)): ~113 neighbours per GP, two 64-lane chunks per GP."""
    return _build("straight21_rho6")


@pytest.fixture(scope="module")
def wider():
    """Support 8.5h: up to ~227 neighbours per GP, four 64-lane chunks."""
    return _build("straight21_rho8.5")


@pytest.mark.parametrize("case", ["c1", "arc", "wide", "wider", "arc_bs"])
def test_neighbours_bit_exact(case, request):
    sh, orc, gpu = request.getfixturevalue(case)
    for d in range(len(gpu.domain_sizes)):
        o_off, o_ids, o_s = orc.neighbours(d, sh.rho)
        g_off, g_ids, g_s = gpu.get_neighbours(d)
        assert np.array_equal(o_off, g_off), f"domain {d}: offsets differ"
        assert np.array_equal(o_ids, g_ids), f"domain {d}: neighbour ids/order differ"
        assert np.array_equal(o_s.view(np.int64), g_s.view(np.int64)), f"domain {d}: s not bit-identical"


@pytest.mark.parametrize("case", ["c1", "arc", "wide", "c1_bs", "arc_bs"])
def test_shape_functions(case, request):
    sh, orc, gpu = request.getfixturevalue(case)
    for d in range(len(gpu.domain_sizes)):
        off, ids, _ = orc.neighbours(d, sh.rho)
        st, _, _ = orc.shape_functions(d, sh.rho, sh.order, sh.kernel)
        assert st == 0
        ref = orc.get_shape_functions(d, off[-1])
        got = gpu.get_shape_functions(d)
        g = pair_gp(off)
        ng = off.size - 1
        e0 = normwise([got[0]], [ref[0]], g, ng)
        e1 = normwise(got[1:3], ref[1:3], g, ng)
        e2 = normwise(got[3:6], ref[3:6], g, ng)
        if d == 0:
            cond = moment_cond(sh.face.gx, sh.face.gy, sh.X, sh.Y, sh.rho, off, ids, ref[0], sh.order, sh.kernel)
        else:
            cond = np.full(ng, 1e4)
        for e, tol, nm in zip((e0, e1, e2), sf_tolerance(cond, SF_TOL), ("psi", "grad", "hess")):
            frac = float(np.mean(e <= SF_TOL))
            print(f"domain {d} {nm}: max {e.max():.3e}, median {np.median(e):.3e}, within 1e-12: {frac:.4f}")
            bad = np.nonzero(e > tol)[0]
            assert bad.size == 0, f"domain {d} {nm}: {bad.size} GPs above tolerance, worst {e.max():.3e}"
            if d == 0:
                assert frac >= FRAC_MIN[nm], f"{nm}: only {frac:.4f} of the face GPs within 1e-12"


@pytest.mark.parametrize("case", ["c1", "arc", "wide", "wider"])
def test_patterns_bit_exact(case, request):
    sh, orc, gpu = request.getfixturevalue(case)
    fo, fi = orc.face_pattern()
    go, gi = gpu.pattern(with_folds=False)
    assert np.array_equal(fo, go) and np.array_equal(fi, gi), "face pattern differs"
    x = W.displacement(sh.nn, 1, 1e-3)
    uo, ui, _, _ = orc.assemble(x, flag_rig=False)
    go, gi = gpu.pattern(with_folds=True)
    assert np.array_equal(uo, go) and np.array_equal(ui, gi), "face+fold union pattern differs"
    # csrMapping of a few GPs
    for gp in (0, 7, sh.face.n // 2, sh.face.n - 1):
        n = int(np.diff(orc.neighbours(0, sh.rho)[0])[gp])
        assert np.array_equal(orc.gp_map(gp, n), gpu.csr_mapping(gp))


@pytest.mark.parametrize("case", ["c1", "arc", "wide", "wider", "c1_bs", "arc_bs"])
@pytest.mark.parametrize("elastic", [True, False])
def test_residual_and_tangent(case, elastic, request):
    sh, orc, gpu = request.getfixturevalue(case)
    x = W.displacement(sh.nn, 1, 1e-3)
    _, _, Kr, Fr = orc.assemble(x, flag_rig=not elastic)
    Fg, Kg = gpu.assemble(x, elastic_folds=elastic)
    ef = np.abs(Fg - Fr).max() / np.abs(Fr).max()
    ek = np.abs(Kg - Kr).max() / np.abs(Kr).max()
    print(f"{case} elastic={elastic}: F rel {ef:.3e}, K rel {ek:.3e}")
    assert ef <= 1e-10
    assert ek <= 1e-12


# Laminates beyond the isotropic ply of the other cases (which has B = 0 and no
# 16/26 couplings in A and D):
# - "coupled": an unsymmetric orthotropic stack (0 / 45 / 90 degrees, unequal
#   thicknesses), A, B and D full, membrane and bending coupled (B != 0);
# - "angle": one orthotropic ply at 30 degrees with a power-of-two thickness,
#   so that B = 0 exactly while A and D have 16/26 terms.
LAMINATES = {
    "coupled": np.array([[1.0, 0.4, 0.25, 0.2, 0.0, 0.004],
                         [1.0, 0.4, 0.25, 0.2, 45.0, 0.003],
                         [1.0, 0.4, 0.25, 0.2, 90.0, 0.005]]),
    "angle": np.array([[1.0, 0.4, 0.25, 0.2, 30.0, 0.0078125]]),
}


@pytest.mark.parametrize("lam", ["coupled", "angle"])
@pytest.mark.parametrize("name", ["straight21", "straight21_rho6"])
def test_laminates(name, lam):
    """The kernels' laminate instantiations other than the isotropic one, with
    up to 64 and with more than 64 neighbours per GP: the device laminate equals
    the oracle's, the residual and tangent meet the bars above, and the FP32
    tangent its 1e-5 bar."""
    sh, orc, gpu = _build(name)
    plies = LAMINATES[lam]
    orc.set_plies(plies)
    gpu.set_plies(plies)
    for a, b in zip(gpu.abd(), orc.abd()):
        assert np.abs(a - b).max() <= 1e-15 * np.abs(b).max()
    A, B, D = gpu.abd()
    assert A[0, 2] != 0.0 and D[1, 2] != 0.0
    assert (np.abs(B).max() > 0.0) == (lam == "coupled")
    x = W.displacement(sh.nn, 2, 1e-3)
    for elastic in (True, False):
        _, _, Kr, Fr = orc.assemble(x, flag_rig=not elastic)
        Fg, Kg = gpu.assemble(x, elastic_folds=elastic)
        ef = np.abs(Fg - Fr).max() / np.abs(Fr).max()
        ek = np.abs(Kg - Kr).max() / np.abs(Kr).max()
        print(f"{name} {lam} elastic={elastic}: F rel {ef:.3e}, K rel {ek:.3e}")
        assert ef <= 1e-10
        assert ek <= 1e-12
    _, K32 = gpu.assemble(x, elastic_folds=False, fp32_tangent=True)
    e32 = np.abs(K32 - Kr).max() / np.abs(Kr).max()
    print(f"{name} {lam} FP32 tangent max rel {e32:.3e}")
    assert e32 <= 1e-5


@pytest.mark.parametrize("case", ["c1", "arc", "wide"])
def test_deterministic(case, request):
    """Bit-identical results run to run, in both tangent types: the FP64 tile's
    waves claim their item chunks in whatever order they finish, and the wide
    case runs the larger workgroups, but every block is still summed by one lane
    in GP order."""
    sh, _, gpu = request.getfixturevalue(case)
    x = W.displacement(sh.nn, 5, 1e-3)
    for fp32 in (False, True):
        runs = [gpu.assemble(x, fp32_tangent=fp32) for _ in range(3)]
        for F, K in runs[1:]:
            assert np.array_equal(runs[0][0].view(np.int64), F.view(np.int64))
            assert np.array_equal(runs[0][1].view(np.int64), K.view(np.int64))


def test_residual_only_matches(c1):
    sh, _, gpu = c1
    x = W.displacement(sh.nn, 6, 1e-3)
    F1, _ = gpu.assemble(x, tangent=True)
    F2, K2 = gpu.assemble(x, tangent=False)
    assert K2 is None
    assert np.array_equal(F1.view(np.int64), F2.view(np.int64))


def test_pivot_collapse_reported_like_reference():
    """At rho = 2.5 h the right-hand GPs next to the straight crease see two
    node columns only: the quadratic basis collapses in the Schur bordering
    (GlobalRKPM.h:303-305, :347-359). Count and first index must match."""
    sh = sheet("straight21_rho2.5")
    orc = O.plate_from_sheet(sh, compute=False)
    orc.neighbours(0, sh.rho)
    st, nb, fb = orc.shape_functions(0, sh.rho, sh.order, sh.kernel)
    assert st == 4
    gpu = G.plate_from_sheet(sh, compute=False)
    gpu.neighbours(0, sh.rho)
    with pytest.raises(G.CvmError) as ei:
        gpu.shape_functions(0, sh.order, sh.kernel)
    assert ei.value.code == G.E_PIVOT_COLLAPSE
    assert (ei.value.nbad, ei.value.first) == (nb, fb)
    assert "pivot collapsed" in str(ei.value)


@pytest.mark.parametrize("case", ["c1", "arc", "arc_bs"])
def test_exact_hinge_tangent(case, request):
    """fold_exact_tangent: the FD dihedral-angle Hessian term (Constraints.h:394-416)."""
    sh, orc, gpu = request.getfixturevalue(case)
    if case == "c1":
        x = W.rigid_fold_displacement(sh, 0.5, seed=3, sigma=1e-3)
    else:
        x = W.displacement(sh.nn, 2, 2e-2)
    _, _, Kr, Fr = orc.assemble(x, flag_rig=False, exact_tangent=True)
    _, _, Kr0, _ = orc.assemble(x, flag_rig=False, exact_tangent=False)
    Fg, Kg = gpu.assemble(x, elastic_folds=True, exact_hinge=True)
    hess_part = np.abs(Kr - Kr0).max()
    assert hess_part > 0.0
    ek = np.abs(Kg - Kr).max()
    print(f"{case}: Hessian term max {hess_part:.3e}, |dK| {ek:.3e} (max|K| {np.abs(Kr).max():.3e})")
    assert ek <= 1e-12 * np.abs(Kr).max() + 1e-6 * hess_part
    assert np.abs(Fg - Fr).max() <= 1e-10 * np.abs(Fr).max()


@pytest.mark.parametrize("case", ["c1", "arc", "wide", "c1_bs", "arc_bs"])
def test_fp32_tangent(case, request):
    """Mixed precision: the face tangent in FP32 (inexact-Newton Jacobian), the
    residual still FP64 (equal to the FP64 path up to FMA-contraction round-off
    of the two kernel instantiations). The tangent's error against the FP64
    oracle is stated: <= 1e-5 of max|K|."""
    sh, orc, gpu = request.getfixturevalue(case)
    x = W.displacement(sh.nn, 1, 1e-3)
    _, _, Kr, _ = orc.assemble(x, flag_rig=False)
    F64, _ = gpu.assemble(x, elastic_folds=True)
    F32, K32 = gpu.assemble(x, elastic_folds=True, fp32_tangent=True)
    assert np.abs(F64 - F32).max() <= 1e-13 * np.abs(F64).max()
    ek = np.abs(K32 - Kr).max() / np.abs(Kr).max()
    ef = np.linalg.norm(K32 - Kr) / np.linalg.norm(Kr)
    print(f"{case}: FP32 tangent max rel {ek:.3e}, Frobenius rel {ef:.3e}")
    assert ek <= 1e-5 and ef <= 1e-5


def test_shape_functions_all_matches_per_domain(c1):
    """cvm_shape_functions_all (one status read-back) gives the per-domain results."""
    sh, _, gpu = c1
    ref = [gpu.get_shape_functions(d) for d in range(len(gpu.domain_sizes))]
    gpu.shape_functions_all(sh.order, sh.kernel)
    for d in range(len(gpu.domain_sizes)):
        got = gpu.get_shape_functions(d)
        for q in range(6):
            assert np.array_equal(got[q].view(np.int64), ref[d][q].view(np.int64))


def test_shape_functions_all_reports_failing_domain():
    sh = sheet("straight21_rho2.5")
    gpu = G.plate_from_sheet(sh, compute=False)
    for d in range(len(gpu.domain_sizes)):
        gpu.neighbours(d, sh.rho)
    with pytest.raises(G.CvmError) as ei:
        gpu.shape_functions_all(sh.order, sh.kernel)
    assert ei.value.code == G.E_PIVOT_COLLAPSE and ei.value.domain == 0
    assert "domain 0" in str(ei.value)
    # the same failure through the search and shape functions in one call
    with pytest.raises(G.CvmError) as ei:
        gpu.recompute_shape_functions(sh.order, sh.kernel)
    assert ei.value.code == G.E_PIVOT_COLLAPSE and ei.value.domain == 0
    assert "domain 0" in str(ei.value)


def test_recompute_shape_functions_matches_separate_calls(arc):
    """cvm_recompute_shape_functions (the search and the shape functions of
    every domain in one call, the shape functions formed beside the search's
    verification) gives the lists, shape functions and assembly of
    cvm_refresh_neighbours + cvm_shape_functions_all bit for bit, and reports
    the lists unchanged."""
    sh, _, gpu = arc
    x = W.displacement(sh.nn, 11, 1e-3)
    assert gpu.refresh_neighbours() is False
    gpu.shape_functions_all(sh.order, sh.kernel)
    lists0 = [gpu.get_neighbours(d) for d in range(len(gpu.domain_sizes))]
    sf0 = [gpu.get_shape_functions(d) for d in range(len(gpu.domain_sizes))]
    F0, K0 = gpu.assemble(x)
    for _ in range(2):
        assert gpu.recompute_shape_functions(sh.order, sh.kernel) is False
        for d in range(len(gpu.domain_sizes)):
            for a, b in zip(gpu.get_neighbours(d), lists0[d]):
                assert np.array_equal(a.view(np.int64) if a.dtype == np.float64 else a,
                                      b.view(np.int64) if b.dtype == np.float64 else b), f"domain {d}: lists"
            for a, b in zip(gpu.get_shape_functions(d), sf0[d]):
                assert np.array_equal(a.view(np.int64), b.view(np.int64)), f"domain {d}: shape functions"
        F1, K1 = gpu.assemble(x)
        assert np.array_equal(F1.view(np.int64), F0.view(np.int64))
        assert np.array_equal(K1.view(np.int64), K0.view(np.int64))


def test_refresh_neighbours_unchanged(arc):
    """cvm_refresh_neighbours on unchanged geometry: lists rewritten in place and
    reported unchanged, the plan stays valid, and the shape functions (with
    indicators and node sides evaluated again) and the assembly that follow
    are bit-identical to those before."""
    sh, _, gpu = arc
    x = W.displacement(sh.nn, 9, 1e-3)
    lists0 = [gpu.get_neighbours(d) for d in range(len(gpu.domain_sizes))]
    sf0 = [gpu.get_shape_functions(d) for d in range(len(gpu.domain_sizes))]
    F0, K0 = gpu.assemble(x)
    assert gpu.refresh_neighbours() is False
    gpu.shape_functions_all(sh.order, sh.kernel)
    for d in range(len(gpu.domain_sizes)):
        for a, b in zip(gpu.get_neighbours(d), lists0[d]):
            assert np.array_equal(a.view(np.int64) if a.dtype == np.float64 else a,
                                  b.view(np.int64) if b.dtype == np.float64 else b), f"domain {d}: lists"
        for a, b in zip(gpu.get_shape_functions(d), sf0[d]):
            assert np.array_equal(a.view(np.int64), b.view(np.int64)), f"domain {d}: shape functions"
    F1, K1 = gpu.assemble(x)
    assert np.array_equal(F1.view(np.int64), F0.view(np.int64))
    assert np.array_equal(K1.view(np.int64), K0.view(np.int64))
