"""Batched contexts: many sheets in one context, one launch per stage. Every
sheet's neighbour lists, shape functions, residual and tangent must be
bit-identical to those of a context holding that sheet alone."""
import numpy as np
import pytest

import curvami_gpu as G
import workloads as W
from _cases import sheet

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch():
    sheets = [sheet("straight21"), W.config_batched_sheet(1), sheet("arc41"), W.config_batched_sheet(2),
              W.config_batched_sheet(7)]
    alone = [G.plate_from_sheet(sh) for sh in sheets]
    bp = G.batch_from_sheets(sheets)
    return sheets, alone, bp


def test_neighbours_and_shape_functions(batch):
    sheets, alone, bp = batch
    off, ids, s = bp.get_neighbours(bp.face_dom)
    psi_b = bp.get_shape_functions(bp.face_dom)
    for k, (sh, p) in enumerate(zip(sheets, alone)):
        g0, ng = bp.face_range[k]
        o1, i1, s1 = p.get_neighbours(p.face_dom)
        a, b = off[g0], off[g0 + ng]
        assert np.array_equal(off[g0:g0 + ng + 1] - a, o1), f"sheet {k}: offsets"
        assert np.array_equal(ids[a:b] - bp.node0[k], i1), f"sheet {k}: neighbour ids"
        assert np.array_equal(s[a:b].view(np.int64), s1.view(np.int64)), f"sheet {k}: s"
        sf1 = p.get_shape_functions(p.face_dom)
        for q in range(6):
            assert np.array_equal(psi_b[q][a:b].view(np.int64), sf1[q].view(np.int64)), f"sheet {k}: plane {q}"


@pytest.mark.parametrize("fp32", [False, True])
def test_assembly_per_sheet_bit_identical(batch, fp32):
    sheets, alone, bp = batch
    xs = [W.displacement(sh.nn, 11 + k, 1e-3) for k, sh in enumerate(sheets)]
    Fb, Kb = bp.assemble(np.concatenate(xs), fp32_tangent=fp32)
    outer, inner = bp.pattern(with_folds=True)
    for k, (sh, p) in enumerate(zip(sheets, alone)):
        F1, K1 = p.assemble(xs[k], fp32_tangent=fp32)
        n0, nn = bp.node0[k], sh.nn
        assert np.array_equal(Fb[3 * n0:3 * (n0 + nn)].view(np.int64), F1.view(np.int64)), f"sheet {k}: F"
        o1, i1 = p.pattern(with_folds=True)
        seg = outer[3 * n0:3 * (n0 + nn) + 1]
        assert np.array_equal(seg - seg[0], o1), f"sheet {k}: column pointers"
        assert np.array_equal(inner[seg[0]:seg[-1]] - 3 * n0, i1), f"sheet {k}: row indices"
        assert np.array_equal(Kb[seg[0]:seg[-1]].view(np.int64), K1.view(np.int64)), f"sheet {k}: K values"


def test_refresh_neighbours_batched(batch):
    """The device re-search of a batched context reports its lists unchanged and
    leaves neighbours, shape functions and the assembly bit-identical."""
    sheets, _, bp = batch
    xs = np.concatenate([W.displacement(sh.nn, 40 + k, 1e-3) for k, sh in enumerate(sheets)])
    off0, ids0, s0 = bp.get_neighbours(bp.face_dom)
    F0, K0 = bp.assemble(xs)
    assert bp.refresh_neighbours() is False
    bp.shape_functions_all(sheets[0].order, sheets[0].kernel)
    off1, ids1, s1 = bp.get_neighbours(bp.face_dom)
    assert np.array_equal(off0, off1) and np.array_equal(ids0, ids1)
    assert np.array_equal(s0.view(np.int64), s1.view(np.int64))
    F1, K1 = bp.assemble(xs)
    assert np.array_equal(F1.view(np.int64), F0.view(np.int64))
    assert np.array_equal(K1.view(np.int64), K0.view(np.int64))
    # the search and the shape functions in one call: the same bits
    sf1 = bp.get_shape_functions(bp.face_dom)
    assert bp.recompute_shape_functions(sheets[0].order, sheets[0].kernel) is False
    for a, b in zip(bp.get_shape_functions(bp.face_dom), sf1):
        assert np.array_equal(a.view(np.int64), b.view(np.int64))
    F2, K2 = bp.assemble(xs)
    assert np.array_equal(F2.view(np.int64), F0.view(np.int64))
    assert np.array_equal(K2.view(np.int64), K0.view(np.int64))
