"""The reference's own inputs end to end on the device: sheets built from its
meshes and configurations (mesh_setup: Gmsh reader, 3 x 3 quad rule,
1-point line rule, creases from cuts / fold tags, rho = 6 x median 3rd-NN
spacing, cubic basis as in main.cpp:53, 78) and compared with the oracle.

Bars: neighbour lists and CSC patterns bit-exact; shape functions normwise
per GP to max(1e-12, C eps cond(M)) (C = 4 / 16 / 64 for psi / gradient /
Hessian, M the GP's moment matrix over its visible neighbours: cond is ~2e3 at
the median GP and reaches 1e5-1e6 next to crease crossings and tips);
F_int and K to max(1e-10, 64 eps max cond(M)) relative (inf-norm: an entry
can be dominated by the worst-conditioned GP's Hessian round-off, max over
all domains); constraint triplet pattern bit-exact, values to the same bar
times their scale; external load to 1e-11.
"""
import os

import numpy as np
import pytest

import curvami_gpu as G
import mesh_setup as MS
import oracle as O
import workloads as W
from _cases import moment_cond, normwise, pair_gp, sf_tolerance

pytestmark = pytest.mark.gpu

INPUTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_inputs")
CASES = [("smoke", "CB_n20"), ("al_arccrease", "Arc_LR0p4_Phi30"), ("ls_miura", "Miura"),
         ("tdal_squaretwist", "SquareTwist")]


@pytest.fixture(scope="module", params=CASES, ids=[c for c, _ in CASES])
def built(request):
    cfg, msh = request.param
    print(f"building {cfg} on {msh}.msh", flush=True)
    ms = MS.sheet_from_config(os.path.join(INPUTS, cfg + ".json"), mesh_file=os.path.join(INPUTS, msh + ".msh"))
    sh = ms.sheet
    orc = O.plate_from_sheet(sh, nthreads=8)
    conds = []
    for d in range(len(orc.domains)):
        off, ids, _ = orc.neighbours(d, sh.rho)
        psi = orc.get_shape_functions(d, off[-1])[0]
        gx, gy = _domain_gps(sh, d)
        conds.append(moment_cond(gx, gy, sh.X, sh.Y, sh.rho, off, ids, psi, sh.order))
    return cfg, sh, orc, G.plate_from_sheet(sh), conds


def _value_bar(conds):
    """Bar for sums of shape-function products (F, K, constraint rows): the
    largest per-GP Hessian bar over all domains, at least 1e-10."""
    return max(1e-10, max(sf_tolerance(c.max())[2] for c in conds))


def test_neighbours_and_patterns_bit_exact(built):
    cfg, sh, orc, gpu, _ = built
    for d in range(len(gpu.domain_sizes)):
        o_off, o_ids, o_s = orc.neighbours(d, sh.rho)
        g_off, g_ids, g_s = gpu.get_neighbours(d)
        assert np.array_equal(o_off, g_off) and np.array_equal(o_ids, g_ids), f"{cfg} domain {d}"
        assert np.array_equal(o_s.view(np.int64), g_s.view(np.int64)), f"{cfg} domain {d}: s"
    fo, fi = orc.face_pattern()
    go, gi = gpu.pattern(with_folds=False)
    assert np.array_equal(fo, go) and np.array_equal(fi, gi)
    uo, ui, _, _ = orc.assemble(W.displacement(sh.nn, 1, 1e-3), flag_rig=False)
    go, gi = gpu.pattern(with_folds=True)
    assert np.array_equal(uo, go) and np.array_equal(ui, gi)


def _domain_gps(sh, d):
    """GP coordinates of device domain d in plate_from_sheet's order: face,
    fold left / right pairs, then the boundary domains."""
    doms = [sh.face] + [s for f in sh.folds for s in (f.left, f.right)] + [b.gps for b in sh.boundaries]
    return doms[d].gx, doms[d].gy


def test_shape_functions(built):
    cfg, sh, orc, gpu, conds = built
    for d in range(len(gpu.domain_sizes)):
        off, ids, _ = orc.neighbours(d, sh.rho)
        ref = orc.get_shape_functions(d, off[-1])
        got = gpu.get_shape_functions(d)
        g = pair_gp(off)
        ng = off.size - 1
        tols = sf_tolerance(conds[d])
        for (sl, nm), tol in zip(((slice(0, 1), "psi"), (slice(1, 3), "grad"), (slice(3, 6), "hess")), tols):
            e = normwise(got[sl], ref[sl], g, ng)
            if d == 0:
                print(f"{cfg} face {nm}: max {e.max():.3e}, median {np.median(e):.3e}, "
                      f"worst e/bar {np.max(e / tol):.3f}")
            bad = np.nonzero(e > tol)[0]
            assert bad.size == 0, f"{cfg} domain {d} {nm}: {bad.size} GPs above their bar, worst {e.max():.3e}"


@pytest.mark.parametrize("elastic", [True, False])
def test_residual_and_tangent(built, elastic):
    cfg, sh, orc, gpu, conds = built
    x = W.displacement(sh.nn, 2, 1e-3)
    _, _, Kr, Fr = orc.assemble(x, flag_rig=not elastic)
    Fg, Kg = gpu.assemble(x, elastic_folds=elastic)
    ef = np.abs(Fg - Fr).max() / np.abs(Fr).max()
    ek = np.abs(Kg - Kr).max() / np.abs(Kr).max()
    bar = _value_bar(conds)
    print(f"{cfg} elastic={elastic}: {sh.face.n} face GPs, F rel {ef:.3e}, K rel {ek:.3e} (bar {bar:.2e})")
    assert ef <= bar and ek <= bar


def test_constraints_and_external_force(built):
    cfg, sh, orc, gpu, conds = built
    bar = _value_bar(conds)
    x = W.displacement(sh.nn, 3, 1e-3)
    for rigid in (False, True):
        rows, cols = gpu.constraint_layout(rigid=rigid)
        c_o, r_o, col_o, v_o = orc.constraints(x, flag_rig=rigid)
        assert np.array_equal(rows, r_o) and np.array_equal(cols, col_o), f"{cfg} rigid={rigid}: pattern"
        c_g, v_g = gpu.constraints(x)
        scale_c = max(1e-3, np.abs(c_o).max()) if c_o.size else 1.0
        scale_v = max(1.0, np.abs(v_o).max()) if v_o.size else 1.0
        ec = np.abs(c_g - c_o).max() if c_o.size else 0.0
        ev = np.abs(v_g - v_o).max() if v_o.size else 0.0
        print(f"{cfg} rigid={rigid}: {c_o.size} rows, max|dc| {ec:.2e}, max|dJ| {ev:.2e} (bar {bar:.2e} x scale)")
        assert ec <= bar * scale_c and ev <= bar * scale_v
    Fo = orc.external_force()
    Fg = gpu.external_force()
    if np.abs(Fo).max() == 0.0:
        assert np.abs(Fg).max() == 0.0
    else:
        assert np.abs(Fg - Fo).max() <= 1e-11 * np.abs(Fo).max()
