"""The batched, serving and large-sheet workloads against the oracle and at
full size.

- Batched sheets (256 sheets, sides U{21..121}, one straight crease each): the
  two smallest and two largest sheets of the set, both crease orientations at
  both ends, in one batched context against the oracle of each sheet alone;
  the whole 256-sheet batch in one context, with the FP32 tangent: finite
  results, a rigid motion of every sheet stress-free, repeats bit-identical,
  sampled sheets equal to contexts holding them alone.
- Serving requests (rigid fold of a seeded angle plus noise), one per size
  class up to 201 x 201, through the serving loop against the oracle: whole
  sheets up to 81 x 81, larger sheets on an oracle box sample (every node whose
  support lies in the box has all its face contributions there, so its F_int
  entries and the K entries between two such nodes are exact on the sample).
- The 1001 x 1001 four-vertex sheet (16M face GPs, fold 0 rigid): finite, a
  rigid motion stress-free, K v against central differences of F_int, and the
  oracle box sample around the crease crossing and on the arc.

Bars: F_int 1e-10 of max|F| over the compared entries; K 1e-12 of max|K| over
whole sheets and max(1e-12, 64 eps cond(M)) on box samples that hold GPs next to
a crease (the shape functions' own round-off there), 1e-5 for the FP32 tangent;
tangent consistency 1e-6.
"""
import dataclasses

import numpy as np
import pytest
import scipy.sparse as sp

import curvami_gpu as G
import oracle as O
import workloads as W
from _cases import box_sample_check, k_bar

pytestmark = pytest.mark.gpu

FLAGS = G.RESIDUAL | G.TANGENT | G.ELASTIC_FOLDS
# sheets of the 256-sheet set: (side, vertical crease) = (21, no), (21, yes),
# (121, no), (120, yes): the smallest and largest sides, both orientations
C3_EXTREMES = (194, 39, 81, 188)


def _rigid(X, Y, angle, shift):
    x = np.zeros(3 * X.size)
    c, s = np.cos(angle), np.sin(angle)
    x[0::3] = X * c - Y * s - X + shift[0]
    x[1::3] = X * s + Y * c - Y + shift[1]
    x[2::3] = shift[2]
    return x


def _sym_matvec(outer, inner, K, v):
    """K v for the symmetric K whose upper triangle is stored (selfadjointView<Upper>):
    K = U + U^T - diag(U) - L - L^T, where L holds the strictly lower entries of
    the stored diagonal node blocks, the last three entries of each column."""
    n = outer.size - 1
    U = sp.csc_matrix((K, inner, outer), shape=(n, n))
    y = U @ v + U.T @ v - U.diagonal() * v
    col = np.arange(n)
    node, c = col // 3, col % 3
    for r in range(3):
        m = r > c                                 # row 3 node + r below the diagonal
        pos = outer[col[m] + 1] - 3 + r
        assert np.all(inner[pos] == 3 * node[m] + r)
        lv = K[pos]
        np.subtract.at(y, 3 * node[m] + r, lv * v[col[m]])    # L v
        np.subtract.at(y, col[m], lv * v[3 * node[m] + r])    # L^T v
    return y


# ---------------------------------------------------------------------------
# batched sheets
# ---------------------------------------------------------------------------
def test_batched_extremes_match_oracle():
    sheets = [W.config_batched_sheet(s) for s in C3_EXTREMES]
    sides = [sh.n_side for sh in sheets]
    assert min(sides) == 21 and max(sides) == 121
    bp = G.batch_from_sheets(sheets)
    xs = [W.displacement(sh.nn, 31 + k, 1e-3) for k, sh in enumerate(sheets)]
    Fb, Kb = bp.assemble(np.concatenate(xs))
    _, Kb32 = bp.assemble(np.concatenate(xs), fp32_tangent=True)
    outer, inner = bp.pattern(with_folds=True)
    for k, sh in enumerate(sheets):
        orc = O.plate_from_sheet(sh, nthreads=16)
        uo, ui, Kr, Fr = orc.assemble(xs[k], flag_rig=False, nthreads=16)
        n0, nn = bp.node0[k], sh.nn
        seg = outer[3 * n0:3 * (n0 + nn) + 1]
        assert np.array_equal(seg - seg[0], uo), f"sheet {C3_EXTREMES[k]}: column pointers"
        assert np.array_equal(inner[seg[0]:seg[-1]] - 3 * n0, ui), f"sheet {C3_EXTREMES[k]}: row indices"
        F = Fb[3 * n0:3 * (n0 + nn)]
        K = Kb[seg[0]:seg[-1]]
        K32 = Kb32[seg[0]:seg[-1]]
        ef = np.abs(F - Fr).max() / np.abs(Fr).max()
        ek = np.abs(K - Kr).max() / np.abs(Kr).max()
        e32 = np.abs(K32 - Kr).max() / np.abs(Kr).max()
        print(f"sheet {C3_EXTREMES[k]} (side {sh.n_side}): F rel {ef:.3e}, K rel {ek:.3e}, FP32 K rel {e32:.3e}",
              flush=True)
        assert ef <= 1e-10 and ek <= 1e-12 and e32 <= 1e-5


def test_batch256_full_size():
    """(The context holds ~100 GB of HBM: it is built here and closed at the end.)"""
    sheets = [W.config_batched_sheet(s) for s in range(W.BATCH_SHEETS)]
    print(f"256 sheets: {sum(sh.face.n for sh in sheets)} face GPs", flush=True)
    bp = G.batch_from_sheets(sheets)
    flags = FLAGS | G.FP32_TANGENT
    nn = bp.nn
    x = G.DeviceArray(3 * nn)
    F1, F2 = G.DeviceArray(3 * nn), G.DeviceArray(3 * nn)
    K1, K2 = G.DeviceArray(bp.nnz_union), G.DeviceArray(bp.nnz_union)
    xv = np.concatenate([W.displacement(sh.nn, 500 + k, 1e-3) for k, sh in enumerate(sheets)])
    x.set(xv)
    bp.assemble_device(x.ptr, F1.ptr, K1.ptr, flags)
    bp.assemble_device(x.ptr, F2.ptr, K2.ptr, flags)
    bp.synchronize()
    Fa, Fb = F1.get(), F2.get()
    assert np.all(np.isfinite(Fa))
    assert np.array_equal(Fa.view(np.int64), Fb.view(np.int64)), "repeat: F differs"
    chunk = 1 << 27
    kmax = 0.0
    for lo in range(0, bp.nnz_union, chunk):
        hi = min(lo + chunk, bp.nnz_union)
        a, b = K1.get_range(lo, hi), K2.get_range(lo, hi)
        assert np.all(np.isfinite(a)), f"K not finite in [{lo}, {hi})"
        assert np.array_equal(a.view(np.int64), b.view(np.int64)), f"repeat: K differs in [{lo}, {hi})"
        kmax = max(kmax, float(np.abs(a).max()))
    print(f"batch: |F| {np.abs(Fa).max():.3e}, |K| {kmax:.3e}, repeats bit-identical", flush=True)
    # sampled sheets against contexts holding them alone
    outer, _ = bp.pattern(with_folds=True)
    for k in (0, 81, 194, 255):
        sh = sheets[k]
        p = G.plate_from_sheet(sh)
        Fs, Ks = p.assemble(xv[3 * bp.node0[k]:3 * (bp.node0[k] + sh.nn)], fp32_tangent=True)
        n0 = bp.node0[k]
        assert np.array_equal(Fa[3 * n0:3 * (n0 + sh.nn)].view(np.int64), Fs.view(np.int64)), f"sheet {k}: F"
        seg = outer[3 * n0:3 * (n0 + sh.nn) + 1]
        assert np.array_equal(K1.get_range(seg[0], seg[-1]).view(np.int64), Ks.view(np.int64)), f"sheet {k}: K"
        p.close()
    # a rigid motion of every sheet (its own angle and shift) is stress-free
    rng = np.random.default_rng(3)
    xr = np.concatenate([_rigid(sh.X, sh.Y, rng.uniform(-1, 1), rng.uniform(-0.2, 0.2, 3)) for sh in sheets])
    x.set(xr)
    bp.assemble_device(x.ptr, F1.ptr, 0, G.RESIDUAL | G.ELASTIC_FOLDS)
    bp.synchronize()
    Fr = F1.get()
    print(f"rigid motion: |F| {np.abs(Fr).max():.3e} against {np.abs(Fa).max():.3e}", flush=True)
    assert np.abs(Fr).max() <= 1e-9 * np.abs(Fa).max()
    for a in (x, F1, F2, K1, K2):
        a.free()
    bp.close()


# ---------------------------------------------------------------------------
# serving requests
# ---------------------------------------------------------------------------
# (side, oracle box or None for the whole sheet)
SERVE_CLASSES = [(21, None), (81, None), (141, (0.4, 0.6, 0.4, 0.6)), (201, (0.4, 0.6, 0.3, 0.5))]


@pytest.mark.parametrize("fp32", [False, True])
def test_serving_size_classes_match_oracle(fp32):
    flags = FLAGS | (G.FP32_TANGENT if fp32 else 0)
    sides = [n for n, _ in SERVE_CLASSES]
    sheets = {n: W.serving_sheet(n) for n in sides}
    plates = {n: G.plate_from_sheet(sheets[n]) for n in sides}
    srv = G.Server(hbm_budget_bytes=32 << 30, max_inflight=4)
    sid = {n: srv.add_sheet(plates[n]) for n in sides}
    reqs = []
    for k, (n, _) in enumerate(SERVE_CLASSES):
        x = W.rigid_fold_displacement(sheets[n], 0.4 + 0.5 * k, seed=900 + k, sigma=1e-3)
        F = np.empty(3 * sheets[n].nn)
        K = G.DeviceArray(plates[n].nnz_union)
        reqs.append((n, x, F, K, srv.submit(sid[n], x, F, flags, K)))     # all in flight together
    for n, x, F, K, rid in reqs:
        srv.wait(rid)
        Kv = K.get()
        K.free()
        box = dict(SERVE_CLASSES)[n]
        sh = sheets[n]
        if box is None:
            orc = O.plate_from_sheet(sh, nthreads=16)
            uo, ui, Kr, Fr = orc.assemble(x, flag_rig=False, nthreads=16)
            outer, inner = plates[n].pattern(with_folds=True)
            assert np.array_equal(outer, uo) and np.array_equal(inner, ui)
            ef = np.abs(F - Fr).max() / np.abs(Fr).max()
            ek = np.abs(Kv - Kr).max() / np.abs(Kr).max()
            print(f"request side {n}: F rel {ef:.3e}, K rel {ek:.3e}", flush=True)
            kbar = k_bar(0.0, fp32)
        else:
            outer, inner = plates[n].pattern(with_folds=True)
            ef, ek, _, cmax = box_sample_check(sh, x, F, Kv, outer, inner, box, label=f"request side {n}")
            kbar = k_bar(cmax, fp32)
        assert ef <= 1e-10 and ek <= kbar, f"side {n}: F {ef:.3e}, K {ek:.3e} (bar {kbar:.3e})"
    srv.close()


# ---------------------------------------------------------------------------
# the 1001 x 1001 four-vertex sheet
# ---------------------------------------------------------------------------
def test_large_sheet_full_size():
    """(About 200 GB of HBM with the FP64 scratch: built here, closed at the end.)"""
    import time
    t0 = time.perf_counter()

    def tick(msg):
        print(f"[{time.perf_counter() - t0:6.1f} s] {msg}", flush=True)
    sh = W.config5(1001)
    sh = dataclasses.replace(sh, folds=[dataclasses.replace(sh.folds[0], rigid=True)] + list(sh.folds[1:]))
    tick(f"{sh.name}: {sh.nn} nodes, {sh.face.n} face GPs")
    p = G.plate_from_sheet(sh)
    tick("device plate set up")
    # noise of 0.2 node spacings, as 1e-3 is on the 201 x 201 sheet: strains of
    # ~0.2, where central differences with step 1e-6 are accurate to ~1e-8
    hx = 1.0 / (sh.n_side - 1)
    x = W.displacement(sh.nn, 1, 0.2 * hx)
    F, K = p.assemble(x)
    assert np.all(np.isfinite(F)) and np.all(np.isfinite(K))
    outer, inner = p.pattern(with_folds=True)
    tick(f"K: {K.size} values, max {np.abs(K).max():.3e}; |F| {np.abs(F).max():.3e}")
    # the oracle box sample around the crossing of the straight creases, and on
    # the arc where the vertical crease meets it
    for box in ((0.47, 0.53, 0.47, 0.53), (0.47, 0.53, 0.77, 0.83)):
        ef, ek, _, cmax = box_sample_check(sh, x, F, K, outer, inner, box, label="1001x1001")
        assert ef <= 1e-10 and ek <= k_bar(cmax), f"box {box}: F {ef:.3e}, K {ek:.3e}"
        tick("box checked")
    del K
    # tangent consistent with central differences of the residual (the exact,
    # FD-Hessian hinge tangent: the default hinge tangent is rank-one)
    F, K = p.assemble(x, exact_hinge=True)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(x.size)
    v /= np.linalg.norm(v)
    h = 1e-6 * max(1.0, np.abs(x).max())
    Fp, _ = p.assemble(x + h * v, tangent=False, exact_hinge=True)
    Fm, _ = p.assemble(x - h * v, tangent=False, exact_hinge=True)
    fd = (Fp - Fm) / (2 * h)
    del Fp, Fm
    tick("central differences")
    Kv = _sym_matvec(outer, inner, K, v)
    rel = np.linalg.norm(Kv - fd) / np.linalg.norm(fd)
    tick(f"|K v - FD| / |FD| = {rel:.3e}")
    assert rel < 1e-6
    del K, Kv
    # rigid motion
    Fr, _ = p.assemble(_rigid(sh.X, sh.Y, 0.3, (0.2, -0.1, 0.05)), tangent=False)
    tick(f"rigid motion: |F| {np.abs(Fr).max():.3e} against {np.abs(F).max():.3e}")
    assert np.abs(Fr).max() <= 1e-9 * np.abs(F).max()
    p.close()
