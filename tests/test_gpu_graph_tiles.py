"""Captured assembly graph, persistent tile grid and tile shapes."""
import numpy as np
import pytest

import autotune as AT
import curvami_gpu as G
import workloads as W
from _cases import sheet

pytestmark = pytest.mark.gpu

FLAGS = G.RESIDUAL | G.TANGENT | G.ELASTIC_FOLDS


@pytest.fixture(scope="module")
def arc():
    sh = sheet("arc41")
    return sh, G.plate_from_sheet(sh)


def _dev(p, sh):
    x = G.DeviceArray(3 * sh.nn); F = G.DeviceArray(3 * sh.nn); K = G.DeviceArray(p.nnz_union)
    return x, F, K


def test_graph_replays_match_eager(arc):
    """Replays with the displacement updated in place between them, eager
    evaluations and host synchronisations interleaved; every replay equals the
    eager result bit for bit."""
    sh, p = arc
    x, F, K = _dev(p, sh)
    x.set(W.displacement(sh.nn, 1, 1e-3))
    for fp32 in (False, True):
        flags = FLAGS | (G.FP32_TANGENT if fp32 else 0)
        p.graph_capture(x.ptr, F.ptr, K.ptr, flags)
        for rep in range(4):
            xv = W.displacement(sh.nn, 20 + rep, 1e-3 * (rep + 1))
            x.set(xv)
            p.graph_launch()
            p.synchronize()
            Fg, Kg = F.get(), K.get()
            Fe, Ke = p.assemble(xv, fp32_tangent=fp32)          # eager, between replays
            assert np.array_equal(Fg.view(np.int64), Fe.view(np.int64)), f"replay {rep}: F"
            assert np.array_equal(Kg.view(np.int64), Ke.view(np.int64)), f"replay {rep}: K"


def test_persistent_grid_bit_identical(arc):
    sh, p = arc
    xv = W.displacement(sh.nn, 3, 1e-3)
    F0, K0 = p.assemble(xv)
    for grid in (1, 7, 256):
        p.set_tile_shape(16, 0, grid)
        F1, K1 = p.assemble(xv)
        assert np.array_equal(F1.view(np.int64), F0.view(np.int64))
        assert np.array_equal(K1.view(np.int64), K0.view(np.int64))
    p.set_tile_shape(16, 0, 0)


def test_tile_shapes_agree_to_roundoff(arc):
    sh, p = arc
    xv = W.displacement(sh.nn, 4, 1e-3)
    F0, K0 = p.assemble(xv)
    for tg, um in ((8, 64), (12, 48), (4, 32)):
        p.set_tile_shape(tg, um, 0)
        p.init_assembly()
        shape = p.tile_shape()
        assert shape["tile_gps"] <= tg
        F1, K1 = p.assemble(xv)
        assert np.abs(F1 - F0).max() <= 1e-13 * np.abs(F0).max()
        assert np.abs(K1 - K0).max() <= 1e-13 * np.abs(K0).max()
    p.set_tile_shape(16, 0, 0)
    p.init_assembly()


def test_autotune_picks_a_candidate(arc):
    sh, p = arc
    x, F, K = _dev(p, sh)
    x.set(W.displacement(sh.nn, 5, 1e-3))
    res = AT.autotune(p, x.ptr, F.ptr, K.ptr, FLAGS, tile_gps=(16, 8), union_max=(64,), grids=(0, 512), reps=2)
    assert len(res["table"]) == 4 and res["ms"] == min(r["ms"] for r in res["table"])
    tg, um, gr = res["best"]
    assert p.tile_shape()["grid"] == gr
    p.set_tile_shape(16, 0, 0)
    p.init_assembly()


def _refused(p):
    with pytest.raises(G.CvmError) as e:
        p.graph_launch()
    assert e.value.code == G.E_STATE, str(e.value)


def test_graph_refused_after_changes(arc):
    """A graph replays the launches of its capture. An eager FP64 evaluation
    between replays of an FP32-tangent capture keeps it valid (the capture
    reserved the FP64 scratch, so nothing it uses is reallocated); a changed
    tile grid, laminate or rebuilt plan makes the replay refuse with
    CVM_E_STATE until the graph is captured again."""
    sh, p = arc
    x, F, K = _dev(p, sh)
    xv = W.displacement(sh.nn, 7, 1e-3)
    x.set(xv)
    p.graph_capture(x.ptr, F.ptr, K.ptr, FLAGS | G.FP32_TANGENT)
    p.assemble(xv)                                   # eager FP64 between replays
    p.graph_launch()
    p.synchronize()
    Fe, Ke = p.assemble(xv, fp32_tangent=True)
    assert np.array_equal(F.get().view(np.int64), Fe.view(np.int64))
    assert np.array_equal(K.get().view(np.int64), Ke.view(np.int64))
    # tile grid
    p.set_tile_shape(16, 0, 64)
    _refused(p)
    p.set_tile_shape(16, 0, 0)
    p.graph_launch()                                 # the captured launches again
    p.synchronize()
    # laminate
    plies = sh.plies.copy()
    other = plies.copy()
    other[:, 0] *= 2.0
    other[:, 1] *= 2.0
    p.set_plies(other)
    _refused(p)
    p.set_plies(plies)
    p.graph_launch()
    p.synchronize()
    # plan rebuilt with another tile shape
    p.set_tile_shape(8, 64, 0)
    p.init_assembly()
    _refused(p)
    p.set_tile_shape(16, 0, 0)
    p.init_assembly()
    p.graph_capture(x.ptr, F.ptr, K.ptr, FLAGS)
    p.graph_launch()
    p.synchronize()
    Fe, Ke = p.assemble(xv)
    assert np.array_equal(F.get().view(np.int64), Fe.view(np.int64))
    assert np.array_equal(K.get().view(np.int64), Ke.view(np.int64))
