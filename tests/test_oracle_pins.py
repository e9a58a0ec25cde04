"""Pins of the CPU oracle (oracle/oracle.cpp) by closed forms, invariants and
brute force, chosen so that a dropped term, a wrong sign or index, or a
transposed operand fails at least one of them. The reference ships no golden
vectors for this path (its regression gate compares two live builds), so
these checks are what fixes the oracle.

CPU only: no GPU, no HIP library.
"""
import math

import numpy as np
import pytest

import oracle as O
import workloads as W
from _cases import RHO_FACTOR, pair_gp

EPS = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def c1():
    sh = W.config1(rho_factor=RHO_FACTOR)
    return sh, O.plate_from_sheet(sh)


@pytest.fixture(scope="module")
def arc():
    sh = W.config2(rho_factor=RHO_FACTOR, n_side=41, nseg=64)
    return sh, O.plate_from_sheet(sh)


# ---- a3: window functions (GlobalRKPM.h:15-17) -----------------------------------
def test_quartic_window_closed_form():
    s = np.array([0.0, 1e-15, 0.25, 0.5, 0.75, 1.0 - 1e-12, 1.0, 1.5])
    w, dw, d2w = O.window(0, s)
    assert w[0] == 1.0 and d2w[0] == -12.0 and dw[0] == 0.0 and dw[1] == 0.0
    assert np.allclose(w[2:5], 1 - 6 * s[2:5] ** 2 + 8 * s[2:5] ** 3 - 3 * s[2:5] ** 4, rtol=0, atol=1e-15)
    assert abs(w[5]) < 1e-14 and abs(dw[5]) < 1e-10 and abs(d2w[5]) < 1e-10   # C2 at s = 1
    assert w[6] == dw[6] == d2w[6] == 0.0 and w[7] == 0.0
    # derivatives against central differences
    ss = np.linspace(0.05, 0.95, 19)
    h = 1e-6
    wp, dwp, _ = O.window(0, ss + h)
    wm, dwm, _ = O.window(0, ss - h)
    _, dw0, d2w0 = O.window(0, ss)
    assert np.allclose((wp - wm) / (2 * h), dw0, atol=1e-8)
    assert np.allclose((dwp - dwm) / (2 * h), d2w0, atol=1e-7)


def test_cubic_bspline_window():
    ss = np.concatenate([np.linspace(0.01, 0.99, 99), [0.5 - 1e-9, 0.5 + 1e-9]])
    w, dw, d2w = O.window(1, ss)
    h = 1e-6
    wp, dwp, _ = O.window(1, ss + h)
    wm, dwm, _ = O.window(1, ss - h)
    assert np.allclose((wp - wm) / (2 * h), dw, atol=1e-7)
    assert np.allclose((dwp - dwm) / (2 * h), d2w, atol=1e-5)
    w0, _, d2 = O.window(1, np.array([0.0, 1.0]))
    assert abs(w0[0] - 2.0 / 3.0) < 1e-15 and w0[1] == 0.0 and d2[0] == -8.0
    # C2 continuity at s = 1/2
    assert abs(w[-2] - w[-1]) < 1e-8 and abs(dw[-2] - dw[-1]) < 1e-8 and abs(d2w[-2] - d2w[-1]) < 1e-7


# ---- a4: crease indicator (Enrichment.h:23-92) -------------------------------------
def test_indicator_sides_and_crease():
    p = O.OraclePlate()
    p.add_crease([0.5, 0.5], [0.0, 1.0], True, True)
    assert p.indicator(0, 0.3, 0.4)[0] == 1.0          # left of an upward segment: normal (-1, 0)
    assert p.indicator(0, 0.7, 0.4)[0] == -1.0
    on = p.indicator(0, 0.5, 0.4)
    assert on[0] == 0.0 and on[7] == 0.0               # on the crease: invalid, f = 0
    # extended ends: beyond the polyline the sign still follows the line
    assert p.indicator(0, 0.3, 1.7)[0] == 1.0 and p.indicator(0, 0.7, -0.6)[0] == -1.0


def test_indicator_cap_gradient_hessian():
    """At a free end (no extension) f = u_hat . n with analytic gradient and
    Hessian (Enrichment.h:75-85); check them against finite differences."""
    p = O.OraclePlate()
    p.add_crease([0.2, 0.6], [0.3, 0.3], False, False)
    for (x, y) in [(0.7, 0.35), (0.75, 0.2), (0.1, 0.42)]:
        out = p.indicator(0, x, y)
        f, gx, gy, hxx, hyy, hxy = out[0], out[2], out[3], out[4], out[5], out[6]
        ux, uy = (x - (0.6 if x > 0.6 else 0.2)), (y - 0.3)
        d = math.hypot(ux, uy)
        assert abs(f - uy / d) < 1e-14                  # segment normal is (0, 1)
        h = 1e-6
        fx = (p.indicator(0, x + h, y)[0] - p.indicator(0, x - h, y)[0]) / (2 * h)
        fy = (p.indicator(0, x, y + h)[0] - p.indicator(0, x, y - h)[0]) / (2 * h)
        assert abs(fx - gx) < 1e-7 and abs(fy - gy) < 1e-7
        gxp = p.indicator(0, x + h, y); gxm = p.indicator(0, x - h, y)
        gyp = p.indicator(0, x, y + h); gym = p.indicator(0, x, y - h)
        assert abs((gxp[2] - gxm[2]) / (2 * h) - hxx) < 1e-6
        assert abs((gyp[3] - gym[3]) / (2 * h) - hyy) < 1e-6
        assert abs((gyp[2] - gym[2]) / (2 * h) - hxy) < 1e-6


# ---- a2: neighbour search (FRNN.h:23-78) ---------------------------------------------
@pytest.mark.parametrize("case", ["c1", "arc"])
def test_frnn_brute_force_and_order(case, request):
    sh, p = request.getfixturevalue(case)
    off, ids, s = p.neighbours(0, sh.rho)
    X, Y = sh.X, sh.Y
    mnx, mny = X.min(), Y.min()
    rho = sh.rho
    rho2 = rho * rho
    rng = np.random.default_rng(0)
    for g in rng.choice(sh.face.n, 300, replace=False):
        gx, gy = sh.face.gx[g], sh.face.gy[g]
        ex, ey = X - gx, Y - gy
        d2 = ex * ex + ey * ey
        want = set(np.nonzero(d2 <= rho2)[0].tolist())
        got = ids[off[g]:off[g + 1]]
        assert set(got.tolist()) == want and len(got) == len(want)
        # order: stencil cell dx outer, dy inner, ascending node index within a cell
        qx = math.floor((gx - mnx) / rho); qy = math.floor((gy - mny) / rho)
        key = [((math.floor((X[i] - mnx) / rho) - qx), (math.floor((Y[i] - mny) / rho) - qy), i) for i in got]
        assert key == sorted(key)
        assert np.array_equal(s[off[g]:off[g + 1]], np.sqrt(d2[got]) / rho)


# ---- a5-a7: RKPM shape functions (GlobalRKPM.h:19-397) --------------------------------
@pytest.mark.parametrize("case", ["c1", "arc"])
def test_rkpm_reproducing_conditions(case, request):
    """sum_I Psi_I X_I^a Y_I^b = x^a y^b for a + b <= 2, with first and second
    derivatives, at every face GP (each side of a crease reproduces on its own)."""
    sh, p = request.getfixturevalue(case)
    off, ids, _ = p.neighbours(0, sh.rho)
    psi, dpx, dpy, dxx, dyy, dxy = p.get_shape_functions(0, off[-1])
    g = pair_gp(off); ng = sh.face.n
    x, y = sh.face.gx, sh.face.gy
    X, Y = sh.X[ids], sh.Y[ids]
    S = lambda v: np.bincount(g, v, ng)
    h = 1.0 / (sh.n_side - 1)
    for a, b in [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2)]:
        m = X ** a * Y ** b
        ex = x ** a * y ** b
        ex_x = a * x ** max(a - 1, 0) * y ** b
        ex_y = b * x ** a * y ** max(b - 1, 0)
        ex_xx = a * (a - 1) * x ** max(a - 2, 0) * y ** b
        ex_yy = b * (b - 1) * x ** a * y ** max(b - 2, 0)
        ex_xy = a * b * x ** max(a - 1, 0) * y ** max(b - 1, 0)
        assert np.abs(S(psi * m) - ex).max() < 1e-11
        assert np.abs(S(dpx * m) - ex_x).max() < 1e-9 / h
        assert np.abs(S(dpy * m) - ex_y).max() < 1e-9 / h
        assert np.abs(S(dxx * m) - ex_xx).max() < 1e-7 / h ** 2
        assert np.abs(S(dyy * m) - ex_yy).max() < 1e-7 / h ** 2
        assert np.abs(S(dxy * m) - ex_xy).max() < 1e-7 / h ** 2


def test_rkpm_crease_cut(c1):
    """Nodes across the crease get Psi = grad Psi = Hessian = 0 exactly; nodes on
    the crease belong to the + side (f >= 0, Enrichment/GlobalRKPM.h:93)."""
    sh, p = c1
    off, ids, _ = p.neighbours(0, sh.rho)
    sf = p.get_shape_functions(0, off[-1])
    g = pair_gp(off)
    left_gp = sh.face.gx[g] < 0.5
    node_left = sh.X[ids] <= 0.5 + 1e-12            # + side: x <= 0.5 including the crease column
    cut = left_gp != node_left
    for a in sf:
        assert np.all(a[cut] == 0.0)


def _independent_shape_functions(sh, p, gps, off, ids):
    """Shape functions rebuilt with a different algorithm (dense moment
    matrices from numpy products, Cholesky solves for M^-1 e0 and its
    derivatives), from the same window jets (restated from GlobalRKPM.h:59-112)."""
    rho = sh.rho
    ex = [0, 1, 0, 2, 1, 0]; ey = [0, 0, 1, 0, 1, 2]
    res = [np.zeros(off[-1]) for _ in range(6)]
    for q in gps:
        b, e = off[q], off[q + 1]
        I = ids[b:e]
        xi = (sh.face.gx[q] - sh.X[I]) / rho; eta = (sh.face.gy[q] - sh.Y[I]) / rho
        s = np.sqrt(xi * xi + eta * eta)
        w, d, dd = O.window(sh.kernel, s)
        dx = xi / s; dy = eta / s; s3 = s ** 3
        jet = [w, d * dx, d * dy, dd * dx * dx + d * xi * xi / s3, dd * dy * dy + d * eta * eta / s3,
               dd * dx * dy - d * xi * eta / s3]
        f = p.indicator(0, sh.face.gx[q], sh.face.gy[q])
        sn = np.array([1.0 if p.indicator(0, sh.X[i], sh.Y[i])[0] >= 0 else -1.0 for i in I])
        Wv = 0.5 * (1 + sn * f[0])
        jet = [j * Wv for j in jet]               # straight crease: f = +-1, no gradient
        P = np.array([xi ** ex[k] * eta ** ey[k] for k in range(6)])
        PX = np.array([ex[k] * xi ** max(ex[k] - 1, 0) * eta ** ey[k] for k in range(6)])
        PY = np.array([ey[k] * xi ** ex[k] * eta ** max(ey[k] - 1, 0) for k in range(6)])
        PXX = np.array([ex[k] * (ex[k] - 1) * xi ** max(ex[k] - 2, 0) * eta ** ey[k] for k in range(6)])
        PYY = np.array([ey[k] * (ey[k] - 1) * xi ** ex[k] * eta ** max(ey[k] - 2, 0) for k in range(6)])
        PXY = np.array([ex[k] * ey[k] * xi ** max(ex[k] - 1, 0) * eta ** max(ey[k] - 1, 0) for k in range(6)])
        ph, px_, py_, pxx, pyy, pxy = jet
        M = (P * ph) @ P.T
        Mx = (P * px_) @ P.T + (PX * ph) @ P.T + (P * ph) @ PX.T
        My = (P * py_) @ P.T + (PY * ph) @ P.T + (P * ph) @ PY.T
        Mxx = (P * pxx) @ P.T + 2 * ((PX * px_) @ P.T + (P * px_) @ PX.T) + (PXX * ph) @ P.T + 2 * (PX * ph) @ PX.T + (P * ph) @ PXX.T
        Myy = (P * pyy) @ P.T + 2 * ((PY * py_) @ P.T + (P * py_) @ PY.T) + (PYY * ph) @ P.T + 2 * (PY * ph) @ PY.T + (P * ph) @ PYY.T
        Mxy = ((P * pxy) @ P.T + (PY * px_) @ P.T + (P * px_) @ PY.T + (PX * py_) @ P.T + (P * py_) @ PX.T
               + (PXY * ph) @ P.T + (PX * ph) @ PY.T + (PY * ph) @ PX.T + (P * ph) @ PXY.T)
        L = np.linalg.cholesky(M)
        sv = lambda r: np.linalg.solve(L.T, np.linalg.solve(L, r))
        c = sv(np.eye(6)[0]); cx = -sv(Mx @ c); cy = -sv(My @ c)
        cxx = -sv(Mxx @ c + 2 * Mx @ cx); cyy = -sv(Myy @ c + 2 * My @ cy); cxy = -sv(Mxy @ c + Mx @ cy + My @ cx)
        C0 = c @ P; Cx = cx @ P + c @ PX; Cy = cy @ P + c @ PY
        Cxx = cxx @ P + 2 * cx @ PX + c @ PXX; Cyy = cyy @ P + 2 * cy @ PY + c @ PYY
        Cxy = cxy @ P + cx @ PY + cy @ PX + c @ PXY
        res[0][b:e] = C0 * ph
        res[1][b:e] = (Cx * ph + C0 * px_) / rho
        res[2][b:e] = (Cy * ph + C0 * py_) / rho
        res[3][b:e] = (Cxx * ph + 2 * Cx * px_ + C0 * pxx) / rho ** 2
        res[4][b:e] = (Cyy * ph + 2 * Cy * py_ + C0 * pyy) / rho ** 2
        res[5][b:e] = (Cxy * ph + Cx * py_ + Cy * px_ + C0 * pxy) / rho ** 2
    return res


def test_rkpm_matches_independent_solve(c1):
    """Same maths, different algorithm (and rounding): Cholesky derivative solves
    vs the reference's 3x3 inverse + Schur bordering. GPs away from the crease
    (well-conditioned supports) agree to 1e-12 normwise."""
    sh, p = c1
    off, ids, _ = p.neighbours(0, sh.rho)
    ref = p.get_shape_functions(0, off[-1])
    rng = np.random.default_rng(1)
    gps = [q for q in rng.choice(sh.face.n, 200, replace=False) if abs(sh.face.gx[q] - 0.5) > 0.1]
    alt = _independent_shape_functions(sh, p, gps, off, ids)
    for q in gps:
        b, e = off[q], off[q + 1]
        for lo, hi in ((0, 1), (1, 3), (3, 6)):
            num = max(np.abs(alt[k][b:e] - ref[k][b:e]).max() for k in range(lo, hi))
            den = max(np.abs(ref[k][b:e]).max() for k in range(lo, hi))
            assert num / den < 1e-12, (q, lo, num / den)


def test_rkpm_derivatives_vs_finite_differences():
    """grad Psi and d2Psi/dxdy match central differences of the GP position; the
    diagonal second derivatives follow the reference's window jet, which uses
    xi^2/s^3 where the exact second derivative of s has eta^2/s^3
    (GlobalRKPM.h:79-80): d2Psi/dx2 and d2Psi/dy2 are therefore NOT the
    finite-difference values (off by ~30%), while their sum (the Laplacian) is."""
    sh = W.config1(rho_factor=RHO_FACTOR)
    base = np.array([[0.23, 0.31], [0.71, 0.66]])
    h = 1e-5
    shifts = [(0, 0), (h, 0), (-h, 0), (0, h), (0, -h)]
    pts = np.array([(b0 + dx, b1 + dy) for b0, b1 in base for dx, dy in shifts])
    p = O.OraclePlate(); p.set_nodes(sh.X, sh.Y); p.set_plies(sh.plies)
    p.add_domain(W.GPSet(pts[:, 0].copy(), pts[:, 1].copy(), np.ones(len(pts))), is_face=True)
    off, ids, _ = p.neighbours(0, sh.rho)
    assert p.shape_functions(0, sh.rho, 2, 0)[0] == 0
    sf = p.get_shape_functions(0, off[-1])
    fld = lambda k, q: dict(zip(ids[off[q]:off[q + 1]].tolist(), sf[k][off[q]:off[q + 1]]))
    for bi in range(len(base)):
        q0 = 5 * bi
        nodes = list(fld(0, q0))
        def fd(k, a, b_):
            A, B = fld(k, q0 + a), fld(k, q0 + b_)
            return np.array([(A.get(i, 0.0) - B.get(i, 0.0)) / (2 * h) for i in nodes])
        at = lambda k: np.array([fld(k, q0)[i] for i in nodes])
        gscale = np.abs(at(1)).max(); hscale = np.abs(at(3)).max()
        assert np.abs(fd(0, 1, 2) - at(1)).max() < 1e-6 * gscale
        assert np.abs(fd(0, 3, 4) - at(2)).max() < 1e-6 * gscale
        assert np.abs(fd(1, 3, 4) - at(5)).max() < 1e-6 * hscale           # d/dy of dPsi/dx
        lap_fd = fd(1, 1, 2) + fd(2, 3, 4)
        assert np.abs(lap_fd - (at(3) + at(4))).max() < 1e-6 * hscale
        assert np.abs(fd(1, 1, 2) - at(3)).max() > 0.1 * hscale          # the reference's d2x quirk


# ---- a9-a12: kinematics and assembly ---------------------------------------------------
def _full_upper(outer, inner, val, n):
    import scipy.sparse as sp
    return sp.csc_matrix((val, inner, outer), shape=(n, n))


def _linear_stiffness(sh, p, off, ids, sf):
    """Closed form of the tangent at u = 0: F0 = [e1 e2], b = e3, no stress, so
    K_IJ = W (Bm_i^T A Bm_j + (beta_i . D beta_j) e3 e3^T) with
    Bm_i = [px e1^T; py e2^T; py e1^T + px e2^T] and beta_i = -(hxx, hyy, 2 hxy)."""
    A, B, D = p.abd()
    nn = sh.nn
    Kfull = np.zeros((3 * nn, 3 * nn))
    _, dpx, dpy, dxx, dyy, dxy = sf
    for q in range(sh.face.n):
        b, e = off[q], off[q + 1]
        I = ids[b:e]
        Bm = np.zeros((e - b, 3, 3))
        Bm[:, 0, 0] = dpx[b:e]; Bm[:, 1, 1] = dpy[b:e]; Bm[:, 2, 0] = dpy[b:e]; Bm[:, 2, 1] = dpx[b:e]
        beta = -np.stack([dxx[b:e], dyy[b:e], 2 * dxy[b:e]], 1)
        Kmm = np.einsum("iak,ab,jbl->ijkl", Bm, A, Bm)
        Kbb = np.einsum("ia,ab,jb->ij", beta, D, beta)
        Kmm[:, :, 2, 2] += Kbb
        W_ = sh.face.w[q]
        idx = (3 * I[:, None] + np.arange(3)[None, :]).ravel()
        Kfull[np.ix_(idx, idx)] += W_ * Kmm.transpose(0, 2, 1, 3).reshape(3 * (e - b), 3 * (e - b))
    return Kfull


def test_zero_displacement_closed_form():
    sh = W.config1(rho_factor=RHO_FACTOR, n_side=11)
    p = O.plate_from_sheet(sh)
    off, ids, _ = p.neighbours(0, sh.rho)
    sf = p.get_shape_functions(0, off[-1])
    outer, inner, val, F = p.assemble(np.zeros(3 * sh.nn), flag_rig=True)
    assert np.all(F == 0.0)
    Ku = _full_upper(outer, inner, val, 3 * sh.nn).toarray()
    Kref = np.triu(_linear_stiffness(sh, p, off, ids, sf))
    # block-upper storage: compare the strictly-upper node blocks and full diagonal blocks
    mask = np.zeros_like(Kref, dtype=bool)
    for i in range(sh.nn):
        for j in range(i, sh.nn):
            mask[3 * i:3 * i + 3, 3 * j:3 * j + 3] = True
    Kref_b = np.where(mask, _linear_stiffness(sh, p, off, ids, sf), 0.0)
    assert np.abs(Ku - Kref_b).max() < 1e-13 * np.abs(Kref_b).max()
    del Kref


def _rotation(ax, ay, az):
    cx, sx, cy, sy, cz, sz = math.cos(ax), math.sin(ax), math.cos(ay), math.sin(ay), math.cos(az), math.sin(az)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def test_rigid_motion_is_stress_free(c1):
    """A large rigid rotation + translation of the whole sheet: facet strains and
    the fold angle stay zero, so F_int vanishes to round-off."""
    sh, p = c1
    R = _rotation(0.7, -0.4, 1.1)
    P0 = np.stack([sh.X, sh.Y, np.zeros(sh.nn)], 1)
    x = (P0 @ R.T + np.array([0.3, -0.2, 0.5]) - P0).ravel()
    _, _, val, F = p.assemble(x, flag_rig=False)
    _, _, _, F_lin = p.assemble(W.displacement(sh.nn, 2, 1e-3), flag_rig=False)
    assert np.abs(F).max() < 1e-10 * np.abs(F_lin).max()


def test_pure_bending_closed_form(c1):
    """w = kappa x^2 / 2 is reproduced exactly by the quadratic basis:
    C0_11 = 1 + kappa^2 x^2 and E1_11 = -kappa / sqrt(1 + kappa^2 x^2)."""
    sh, p = c1
    kappa = 0.8
    x = np.zeros(3 * sh.nn)
    x[2::3] = 0.5 * kappa * sh.X ** 2
    kin = p.kinematics(0, x)
    gx = sh.face.gx
    E0xx, E1xx, E1yy = kin[:, 9], kin[:, 12], kin[:, 13]
    assert np.abs(E0xx - 0.5 * (kappa * gx) ** 2).max() < 1e-11
    assert np.abs(E1xx + kappa / np.sqrt(1 + (kappa * gx) ** 2)).max() < 2e-9
    assert np.abs(E1yy).max() < 2e-9


def _tangent_check(p, x, n, rig, exact=False, ndir=3):
    """TangentCheck (main.cpp:651-710): K v (K read as selfadjointView<Upper>)
    against central differences of F_int along random unit directions."""
    import scipy.sparse as sp
    outer, inner, val, _ = p.assemble(x, flag_rig=rig, exact_tangent=exact)
    K = sp.csc_matrix((val, inner, outer), shape=(n, n))
    Kfull = sp.triu(K) + sp.triu(K, 1).T
    rng = np.random.default_rng(4)
    worst = 0.0
    for _ in range(ndir):
        v = rng.standard_normal(n); v /= np.linalg.norm(v)
        h = 1e-6 * max(1.0, np.abs(x).max())
        fd = (p.internal_force(x + h * v, flag_rig=rig) - p.internal_force(x - h * v, flag_rig=rig)) / (2 * h)
        worst = max(worst, np.linalg.norm(Kfull @ v - fd) / np.linalg.norm(fd))
    return worst


def test_tangent_consistent_with_residual(c1):
    """Face terms only (flagRig = true): consistent to < 1e-6 (the reference's
    CONSISTENT verdict) even at large random deformation."""
    sh, p = c1
    for sigma in (1e-3, 5e-2):
        assert _tangent_check(p, W.displacement(sh.nn, 3, sigma), 3 * sh.nn, True) < 1e-6


def test_hinge_tangent_rank_one_and_exact(c1):
    """The default hinge tangent is the rank-one W Kt g g^T term
    (Constraints.h:422-433), inconsistent by O(theta - theta0); with the
    finite-difference Hessian term (exact_tangent, :394-416) the face + fold
    tangent is consistent again."""
    sh, p = c1
    x = W.rigid_fold_displacement(sh, 0.4, seed=3, sigma=1e-3)
    n = 3 * sh.nn
    assert _tangent_check(p, x, n, False, exact=False) > 1e-6
    assert _tangent_check(p, x, n, False, exact=True) < 1e-6


def test_pattern_brute_force(c1):
    sh, p = c1
    off, ids, _ = p.neighbours(0, sh.rho)
    blocks = set()
    for q in range(sh.face.n):
        I = ids[off[q]:off[q + 1]]
        for a in I:
            for b in I:
                if a <= b:
                    blocks.add((int(a), int(b)))
    outer, inner = p.face_pattern()
    cols = {j: sorted(i for (i, jj) in blocks if jj == j) for j in range(sh.nn)}
    exp_inner = []
    exp_outer = [0]
    for j in range(sh.nn):
        for c in range(3):
            for i in cols[j]:
                exp_inner += [3 * i, 3 * i + 1, 3 * i + 2]
            exp_outer.append(len(exp_inner))
    assert np.array_equal(outer, np.array(exp_outer)) and np.array_equal(inner, np.array(exp_inner))


def test_single_and_multi_thread_agree(c1):
    sh, p = c1
    x = W.displacement(sh.nn, 1, 1e-3)
    _, _, v1, F1 = p.assemble(x, nthreads=1)
    _, _, v8, F8 = p.assemble(x, nthreads=4)
    assert np.abs(v1 - v8).max() <= 1e-14 * np.abs(v1).max()
    assert np.abs(F1 - F8).max() <= 1e-14 * np.abs(F1).max()


# ---- a14: hinge (Constraints.h:71-119, :341-436) ------------------------------------------
def test_fold_angle_of_rigid_fold(c1):
    """Fold the + side rigidly by phi about the crease: every crease GP reads
    |theta| = phi (the facets stay flat), for phi up to 170 degrees."""
    sh, p = c1
    l, r = p.fold_doms[0]
    for deg in (10.0, 90.0, 170.0):
        phi = math.radians(deg)
        x = W.rigid_fold_displacement(sh, phi, seed=0, sigma=0.0)
        kl = p.kinematics(l, x); kr = p.kinematics(r, x)
        for q in range(kl.shape[0]):
            F = np.concatenate([kl[q, 0:3], kl[q, 3:6], kr[q, 0:3], kr[q, 3:6]])
            th, _ = O.fold_angle(F, sh.folds[0].left.tx[q], sh.folds[0].left.ty[q])
            assert abs(abs(th) - phi) < 1e-9


def test_fold_angle_gradient_fd():
    rng = np.random.default_rng(5)
    for _ in range(5):
        F = np.concatenate([[1, 0, 0], [0, 1, 0], [1, 0, 0], [0, 1, 0]]) + 0.2 * rng.standard_normal(12)
        th, g = O.fold_angle(F, 0.0, 1.0)
        h = 1e-7
        for k in range(12):
            Fp = F.copy(); Fp[k] += h
            Fm = F.copy(); Fm[k] -= h
            fd = (O.fold_angle(Fp, 0.0, 1.0)[0] - O.fold_angle(Fm, 0.0, 1.0)[0]) / (2 * h)
            assert abs(fd - g[k]) < 1e-7


def test_hinge_force_is_energy_gradient(c1):
    """With the face stiffness removed (zero laminate), F_int is the hinge force
    W Kt (theta - theta0) dtheta/du: check it against finite differences of the
    hinge energy sum W Kt (theta - theta0)^2 / 2."""
    sh, _ = c1
    p = O.plate_from_sheet(sh)
    p.set_plies(np.array([[1e-30, 1e-30, 0.3, 1e-30, 0.0, 1e-30]]))
    l, r = p.fold_doms[0]
    x = W.rigid_fold_displacement(sh, 0.3, seed=7, sigma=1e-3)
    F = p.internal_force(x, flag_rig=False, with_tangent=False)
    fold = sh.folds[0]

    def energy(xx):
        kl = p.kinematics(l, xx); kr = p.kinematics(r, xx)
        E = 0.0
        for q in range(kl.shape[0]):
            Fv = np.concatenate([kl[q, 0:6], kr[q, 0:6]])
            th, _ = O.fold_angle(Fv, fold.left.tx[q], fold.left.ty[q])
            E += fold.left.w[q] * fold.Kt * (th - fold.Ang0) ** 2 / 2
        return E
    rng = np.random.default_rng(6)
    for k in rng.choice(3 * sh.nn, 6, replace=False):
        h = 1e-6
        xp = x.copy(); xp[k] += h
        xm = x.copy(); xm[k] -= h
        fd = (energy(xp) - energy(xm)) / (2 * h)
        assert abs(fd - F[k]) < 1e-6 * np.abs(F).max() + 1e-16


# An unsymmetric orthotropic stack (0 / 45 / 90 degrees, unequal thicknesses):
# A, B and D are all full, so membrane and bending couple (B != 0).
COUPLED_PLIES = np.array([[1.0, 0.4, 0.25, 0.2, 0.0, 0.004],
                          [1.0, 0.4, 0.25, 0.2, 45.0, 0.003],
                          [1.0, 0.4, 0.25, 0.2, 90.0, 0.005]])


def _clt(plies):
    """Classical lamination theory by strain transformation: with engineering
    strains, the local strains of a ply rotated by theta are T e (T below), and
    equal strain energy gives Qbar = T^T Q T; A, B, D are the moments 0, 1, 2
    of Qbar over the thickness, z from -h/2."""
    h = plies[:, 5].sum()
    A = np.zeros((3, 3)); B = np.zeros((3, 3)); D = np.zeros((3, 3))
    z0 = -h / 2
    for E1, E2, nu12, G12, ang, t in plies:
        nu21 = nu12 * E2 / E1
        den = 1 - nu12 * nu21
        Q = np.array([[E1 / den, nu12 * E2 / den, 0], [nu12 * E2 / den, E2 / den, 0], [0, 0, G12]])
        m, n = math.cos(math.radians(ang)), math.sin(math.radians(ang))
        T = np.array([[m * m, n * n, m * n], [n * n, m * m, -m * n], [-2 * m * n, 2 * m * n, m * m - n * n]])
        Qb = T.T @ Q @ T
        z1 = z0 + t
        A += Qb * (z1 - z0); B += Qb * (z1 ** 2 - z0 ** 2) / 2; D += Qb * (z1 ** 3 - z0 ** 3) / 3
        z0 = z1
    return A, B, D


def test_laminate_abd():
    """computeABD (Constitutive.h:10-50) against the strain-transformation
    restatement above for the unsymmetric stack; a symmetric stack has B = 0
    (to round-off); one isotropic ply gives A = Q t, D = Q t^3 / 12."""
    p = O.OraclePlate()
    p.set_nodes(np.zeros(1), np.zeros(1))
    p.set_plies(COUPLED_PLIES)
    for got, ref in zip(p.abd(), _clt(COUPLED_PLIES)):
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    A, B, D = p.abd()
    assert np.abs(B).max() > 1e-4 * np.abs(A).max() * COUPLED_PLIES[:, 5].sum()
    assert abs(A[0, 2]) > 1e-3 * A[0, 0] and abs(D[1, 2]) > 1e-3 * D[1, 1]
    sym = np.vstack([COUPLED_PLIES, COUPLED_PLIES[::-1]])
    p.set_plies(sym)
    A, B, _ = p.abd()
    assert np.abs(B).max() <= 1e-15 * np.abs(A).max()
    E, nu, t = 2.0, 0.3, 0.01
    p.set_plies(np.array([[E, E, nu, E / (2 * (1 + nu)), 30.0, t]]))
    Q = E / (1 - nu * nu) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    A, B, D = p.abd()
    assert np.abs(A - Q * t).max() <= 1e-14 * E
    assert np.abs(B).max() <= 1e-16 * E
    assert np.abs(D - Q * t ** 3 / 12).max() <= 1e-14 * E * t ** 2


def test_coupled_laminate_residual_and_tangent():
    """With B != 0 the oracle's residual is the gradient of its strain energy
    (1/2) W (e0 . N + e1 . M), N = A e0 + B e1, M = B^T e0 + D e1, and its
    tangent is consistent with the residual (< 1e-6)."""
    sh = W.config1(rho_factor=RHO_FACTOR)
    p = O.plate_from_sheet(sh)
    p.set_plies(COUPLED_PLIES)
    x = W.displacement(sh.nn, 5, 2e-3)
    F = p.internal_force(x, flag_rig=True)
    rng = np.random.default_rng(12)
    for _ in range(3):
        v = rng.standard_normal(x.size)
        h = 1e-6
        e = lambda y: sum(p.energies(y)[:2])
        fd = (e(x + h * v) - e(x - h * v)) / (2 * h)
        assert abs(fd - F @ v) <= 1e-7 * np.linalg.norm(F) * np.linalg.norm(v)
    assert _tangent_check(p, x, 3 * sh.nn, True) < 1e-6


def test_box_sample_is_exact_inside():
    """The oracle box sample used for full-size parity (tests/_cases.py
    box_sample_check): assembling only the face GPs inside a box (and every
    crease GP) gives the full sheet's F_int at every node whose support disk
    lies in the box, and the K entries between two such nodes, bit for bit;
    nodes next to the box edge (support partly outside) do differ, so the
    check is not vacuous."""
    from _cases import box_sample_check, sheet
    sh = sheet("arc41")
    orc = O.plate_from_sheet(sh)
    x = W.displacement(sh.nn, 1, 1e-3)
    outer, inner, K, F = orc.assemble(x, flag_rig=False)
    box = (0.2, 0.8, 0.05, 0.5)
    ef, ek, n, _ = box_sample_check(sh, x, F, K, outer, inner, box, nthreads=1)
    assert ef == 0.0 and ek == 0.0 and n > 100
    # a node whose support crosses the box edge misses contributions
    import dataclasses
    m = (sh.face.gx > box[0]) & (sh.face.gx < box[1]) & (sh.face.gy > box[2]) & (sh.face.gy < box[3])
    sub = W.GPSet(sh.face.gx[m], sh.face.gy[m], sh.face.w[m])
    part = O.plate_from_sheet(dataclasses.replace(sh, face=sub))
    _, _, _, Fp = part.assemble(x, flag_rig=False)
    edge = np.nonzero((np.abs(sh.Y - box[2]) < 0.5 * sh.rho) & (sh.X > 0.4) & (sh.X < 0.6))[0]
    d = np.concatenate([3 * edge, 3 * edge + 1, 3 * edge + 2])
    assert np.abs(Fp[d] - F[d]).max() > 1e-6 * np.abs(F).max()


# ---- natural curvature (Kinematics.h:9-45, :114-115; Meshless.h:212) -----------------
def _band(x, x0, width, theta0, sign):
    """kappa_bar(0, 0) of the band, written from its definition: sign theta0 g(X),
    g(X) = exp(-((X - x0) / w)^2) / (sqrt(pi) w), whose integral over X is 1."""
    return sign * theta0 * np.exp(-((x - x0) / width) ** 2) / (math.sqrt(math.pi) * width)


@pytest.fixture(scope="module")
def natural():
    sh = W.config1(rho_factor=RHO_FACTOR, n_side=21)
    band = (0.37, 0.12, math.radians(30.0), -1.0)          # a mountain band off the crease
    p = O.plate_from_sheet(sh)
    p.set_natural_curvature(True, *band)
    return sh, p, band


def test_natural_curvature_energy_at_rest(natural):
    """At u = 0 the geometric E1 vanishes, so E1_el = kappa_bar eX eX and the
    bending energy is sum_g 1/2 W_g D11 kappa_bar(X_g)^2 (D11 = E t^3 / (12 (1 - nu^2))
    for the one isotropic ply); the membrane energy stays zero. The band's
    integral over X is theta0 (g integrates to 1)."""
    sh, p, band = natural
    Em, Eb, Eh = p.energies(np.zeros(3 * sh.nn))
    D11 = sh.E * sh.t ** 3 / (12.0 * (1.0 - sh.nu ** 2))
    kb = _band(sh.face.gx, *band)
    expect = float(np.sum(0.5 * sh.face.w * D11 * kb ** 2))
    assert Eb == pytest.approx(expect, rel=1e-12)
    assert Em == 0.0 and Eh == 0.0
    xs = np.linspace(-2.0, 3.0, 200001)
    assert np.trapezoid(_band(xs, *band), xs) == pytest.approx(band[3] * band[2], rel=1e-9)


def test_natural_curvature_mountain_sign(natural):
    """Mountain (sign -1): kappa_bar < 0, and E1_el = E1 + kappa_bar is relieved by
    w_xx < 0 (E1 = -w_xx to first order): the bending energy falls along w = c x^2/2
    for c < 0 and rises for c > 0, so the sheet ridges towards +z. A valley does
    the opposite."""
    sh, p, band = natural
    def Eb(c):
        x = np.zeros(3 * sh.nn)
        x[2::3] = 0.5 * c * (sh.X - band[0]) ** 2
        return p.energies(x)[1]
    e0 = Eb(0.0)
    assert Eb(-0.05) < e0 < Eb(0.05)
    p.set_natural_curvature(True, band[0], band[1], band[2], +1.0)
    try:
        e0v = Eb(0.0)
        assert e0v == pytest.approx(e0, rel=1e-14)
        assert Eb(0.05) < e0v < Eb(-0.05)
    finally:
        p.set_natural_curvature(True, *band)


def test_natural_curvature_residual_and_tangent(natural):
    """With the band on, F_int is still dE/dx of E_mem + E_ben + E_hin (so the
    residual uses E1_el exactly where the energy does), the face tangent stays
    consistent with it, and switching the band off gives the plain plate back."""
    sh, p, band = natural
    x = W.displacement(sh.nn, 11, 2e-3)
    F = p.internal_force(x, flag_rig=False)
    h = 1e-7
    for seed in (1, 2):
        v = W.displacement(sh.nn, 200 + seed, 1.0)
        fd = (sum(p.energies(x + h * v)) - sum(p.energies(x - h * v))) / (2 * h)
        assert abs(fd - F @ v) <= 1e-8 * np.linalg.norm(F) * np.linalg.norm(v)
    assert _tangent_check(p, x, 3 * sh.nn, True) < 1e-6
    plain = O.plate_from_sheet(sh)
    F0 = plain.internal_force(x, flag_rig=False)
    assert np.abs(F - F0).max() > 1e-4 * np.abs(F0).max()      # the band changes the residual (~6e-4 here)
    p.set_natural_curvature(False)
    try:
        assert np.array_equal(p.internal_force(x, flag_rig=False), F0)
    finally:
        p.set_natural_curvature(True, *band)
