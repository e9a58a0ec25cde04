"""Shared inputs and comparison helpers for the parity tests."""
import functools

import numpy as np

import workloads as W

# Support radius used for parity: 3.75 h (the configs' "40-50 neighbours").
# At 2.5 h the reference's own RKPM solve reports a pivot collapse on the
# straight-crease sheet (tested as an error-path case).
RHO_FACTOR = 3.75


@functools.lru_cache(maxsize=None)
def sheet(name: str):
    if name == "straight21":
        return W.config1(rho_factor=RHO_FACTOR)
    if name == "straight21_rho2.5":
        return W.config1(rho_factor=2.5)
    if name == "straight21_rho6":
        return W.config1(rho_factor=6.0)
    if name == "straight21_rho8.5":
        return W.config1(rho_factor=8.5)
    if name == "straight11":
        return W.config1(rho_factor=RHO_FACTOR, n_side=11)
    if name == "straight21_bspline":
        return W.config1(rho_factor=RHO_FACTOR, kernel=1)
    if name == "arc41":
        return W.config2(rho_factor=RHO_FACTOR, n_side=41, nseg=64)
    if name == "arc41_bspline":
        return W.config2(rho_factor=RHO_FACTOR, n_side=41, nseg=64, kernel=1)
    if name == "arc201_bspline":
        return W.config2(rho_factor=RHO_FACTOR, kernel=1)
    if name == "arc201":
        return W.config2(rho_factor=RHO_FACTOR)
    raise KeyError(name)


def pair_gp(off):
    return np.repeat(np.arange(off.size - 1), np.diff(off))


def normwise(a, r, g, ng):
    """Per-GP max_I |a - r| / max_I |r| over the listed arrays."""
    num = np.zeros(ng); den = np.zeros(ng)
    for x, y in zip(a, r):
        np.maximum.at(num, g, np.abs(x - y))
        np.maximum.at(den, g, np.abs(y))
    return num / np.where(den > 0, den, 1.0)


EPS = np.finfo(np.float64).eps


def window(s, kernel=0):
    """The window on s = |x - X_I| / rho: 0 the quartic spline
    (GlobalRKPM.h:15-17), 1 the cubic B-spline."""
    if kernel == 1:
        inner = 2.0 / 3.0 - 4 * s ** 2 + 4 * s ** 3
        outer = 4.0 / 3.0 - 4 * s + 4 * s ** 2 - 4.0 / 3.0 * s ** 3
        return np.where(s < 1, np.where(s <= 0.5, inner, outer), 0.0)
    return np.where(s < 1, 1 - 6 * s ** 2 + 8 * s ** 3 - 3 * s ** 4, 0.0)


def moment_cond(gx, gy, X, Y, rho, off, ids, psi, order, kernel=0):
    """cond_2 of each GP's moment matrix M = sum_I w(s_I) p_I p_I^T over its
    visible neighbours (psi != 0), with the given window and the basis of
    the given order in xi = (x - X_I) / rho: the conditioning that bounds the
    round-off two correct evaluations of the shape functions can differ by."""
    g = pair_gp(off)
    xi = (gx[g] - X[ids]) / rho
    eta = (gy[g] - Y[ids]) / rho
    s = np.sqrt(xi * xi + eta * eta)
    w = window(s, kernel) * (psi != 0)
    P = [xi ** (d - k) * eta ** k for d in range(order + 1) for k in range(d + 1)]
    ng = off.size - 1
    M = np.zeros((ng, len(P), len(P)))
    for a in range(len(P)):
        for b in range(a, len(P)):
            M[:, a, b] = M[:, b, a] = np.bincount(g, w * P[a] * P[b], ng)
    return np.linalg.cond(M)


def sf_tolerance(cond, floor=1e-12):
    """Per-GP normwise bars for (psi, gradient, Hessian): max(floor, C eps
    cond(M)) with C = 4 / 16 / 64."""
    return tuple(np.maximum(floor, C * EPS * cond) for C in (4, 16, 64))


def upper_csc(outer, inner, K, n):
    """The stored block-upper K as a scipy CSC matrix (n x n)."""
    import scipy.sparse as sp
    return sp.csc_matrix((K, inner, outer), shape=(n, n))


def box_sample_check(sh, x, F, K, outer, inner, box, nthreads=16, label=""):
    # sh: a workloads.Sheet; x the displacement; F, K, outer, inner the device
    # results and pattern (with the elastic folds) of the full sheet
    """Compare a device residual and tangent of a full sheet with the oracle on
    a bounded sample: the oracle assembles only the face and crease GPs inside
    box = (x0, x1, y0, y1), in their order on the sheet, over the nodes within
    rho of the box (renumbered in ascending order) and the two corners of the
    sheet's bounding box, which keep the neighbour search's cell grid and hence
    every neighbour order as on the full sheet. A node whose support disk
    (radius rho) lies inside the box receives every contribution it has on the
    full sheet, so its F_int entries and the K entries between two such nodes
    must agree with the device's full-sheet values. Returns (F rel error, K rel
    error, number of nodes compared, the largest moment-matrix condition number
    over the box's GPs)."""
    import dataclasses
    import oracle as O
    x0, x1, y0, y1 = box
    r = sh.rho * (1.0 + 1e-6)
    keep = (sh.X > x0 - r) & (sh.X < x1 + r) & (sh.Y > y0 - r) & (sh.Y < y1 + r)
    keep[np.argmin(sh.X + sh.Y)] = keep[np.argmax(sh.X + sh.Y)] = True   # grid corners (min, max)
    assert sh.X[np.argmin(sh.X + sh.Y)] == sh.X.min() and sh.Y[np.argmin(sh.X + sh.Y)] == sh.Y.min()
    assert sh.X[np.argmax(sh.X + sh.Y)] == sh.X.max() and sh.Y[np.argmax(sh.X + sh.Y)] == sh.Y.max()
    nodes = np.nonzero(keep)[0]
    newid = np.full(sh.nn, -1)
    newid[nodes] = np.arange(nodes.size)

    def inside(g):
        m = (g.gx > x0) & (g.gx < x1) & (g.gy > y0) & (g.gy < y1)
        return W.GPSet(*(None if a is None else np.asarray(a)[m] for a in (g.gx, g.gy, g.w, g.nx, g.ny, g.tx, g.ty)))
    sub = inside(sh.face)
    folds = [dataclasses.replace(f, left=inside(f.left), right=inside(f.right)) for f in sh.folds]
    sample = dataclasses.replace(sh, X=sh.X[nodes], Y=sh.Y[nodes], face=sub, folds=folds, boundaries=[])
    xs = x.reshape(-1, 3)[nodes].ravel()
    orc = O.plate_from_sheet(sample, nthreads=nthreads)
    o_outer, o_inner, o_K, o_F = orc.assemble(xs, flag_rig=False, nthreads=1)
    o_off, o_ids, _ = orc.neighbours(0, sh.rho)
    psi = orc.get_shape_functions(0, int(o_off[-1]))[0]
    cond = moment_cond(sub.gx, sub.gy, sample.X, sample.Y, sh.rho, o_off, o_ids, psi, sh.order, sh.kernel)
    eps = 1e-9
    inner_nodes = np.nonzero((sh.X - sh.rho > x0 + eps) & (sh.X + sh.rho < x1 - eps) &
                             (sh.Y - sh.rho > y0 + eps) & (sh.Y + sh.rho < y1 - eps))[0]
    assert inner_nodes.size > 0, "box too small: no node has its support inside"
    dofs = (3 * inner_nodes[:, None] + np.arange(3)[None, :]).ravel()
    sdofs = (3 * newid[inner_nodes][:, None] + np.arange(3)[None, :]).ravel()
    Kg = _csc_block(outer, inner, K, dofs)
    Ko = _csc_block(o_outer, o_inner, o_K, sdofs)
    ef = np.abs(F[dofs] - o_F[sdofs]).max() / np.abs(o_F[sdofs]).max()
    ek = np.abs(Kg - Ko).max() / np.abs(Ko).max()
    print(f"{label} oracle box {box}: {sub.n} GPs, {nodes.size} nodes, {inner_nodes.size} compared, "
          f"F rel {ef:.3e}, K rel {ek:.3e}, max cond(M) {cond.max():.3e}", flush=True)
    return ef, ek, inner_nodes.size, float(cond.max())


def _csc_block(outer, inner, K, dofs):
    """Dense block K[dofs, dofs] of a stored CSC (rows and columns in the order
    of dofs), read from the listed columns only."""
    outer = np.asarray(outer, np.int64)
    lo, hi = outer[dofs], outer[dofs + 1]
    cnt = hi - lo
    pos = np.repeat(lo - np.concatenate(([0], np.cumsum(cnt)[:-1])), cnt) + np.arange(cnt.sum())
    col = np.repeat(np.arange(dofs.size), cnt)
    rowmap = np.full(int(max(dofs.max(), np.asarray(inner).max(initial=0))) + 1, -1)
    rowmap[dofs] = np.arange(dofs.size)
    row = rowmap[np.asarray(inner)[pos]]
    m = row >= 0
    out = np.zeros((dofs.size, dofs.size))
    out[row[m], col[m]] = np.asarray(K)[pos[m]]
    return out


def k_bar(cond_max, fp32=False):
    """Bar for K against the oracle: 1e-12 of max|K|, or 64 eps cond(M) where
    the moment matrices are ill-conditioned (GPs next to a crease, whose second
    shape derivatives carry that much round-off into the bending terms); 1e-5
    for the FP32 tangent."""
    return 1e-5 if fp32 else max(1e-12, 64 * EPS * cond_max)
