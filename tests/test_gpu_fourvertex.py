"""Four-vertex-like layout (two crossing straight creases on node lines and an
arc with interior tips), one fold rigid: neighbours, shape functions, residual,
tangent and constraint rows against the oracle."""
import numpy as np
import pytest

import curvami_gpu as G
import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fv():
    sh = W.config5(41)
    orc = O.plate_from_sheet(sh, nthreads=4)
    gpu = G.plate_from_sheet(sh)
    orc.set_fold_rigid(0)
    gpu.set_fold_rigid(0)
    gpu.init_assembly()
    return sh, orc, gpu


def test_neighbours_bit_exact(fv):
    sh, orc, gpu = fv
    for d in range(len(gpu.domain_sizes)):
        o_off, o_ids, o_s = orc.neighbours(d, sh.rho)
        g_off, g_ids, g_s = gpu.get_neighbours(d)
        assert np.array_equal(o_off, g_off) and np.array_equal(o_ids, g_ids)
        assert np.array_equal(o_s.view(np.int64), g_s.view(np.int64))


def test_shape_functions(fv):
    """Normwise per GP against the oracle; the GPs near the crossing and the arc
    tips are the ill-conditioned ones, so the bar is the conditioning-scaled one
    of the parity suite at its largest (64 eps cond, cond <= 1e5 here)."""
    sh, orc, gpu = fv
    for d in range(len(gpu.domain_sizes)):
        off, ids, _ = orc.neighbours(d, sh.rho)
        st, _, _ = orc.shape_functions(d, sh.rho, sh.order, sh.kernel)
        assert st == 0
        ref = orc.get_shape_functions(d, off[-1])
        got = gpu.get_shape_functions(d)
        for q in range(6):
            err = np.abs(got[q] - ref[q]).max() / max(np.abs(ref[q]).max(), 1e-300)
            assert err <= 1e-9, f"domain {d} plane {q}: {err:.2e}"


@pytest.mark.parametrize("exact", [False, True])
def test_residual_tangent_with_rigid_fold(fv, exact):
    """Folds 1 and 2 elastic, fold 0 rigid: its hinge is absent from F and K."""
    sh, orc, gpu = fv
    x = W.displacement(sh.nn, 7, 1e-3)
    uo, ui, Kr, Fr = orc.assemble(x, flag_rig=False, exact_tangent=exact)
    go, gi = gpu.pattern(with_folds=True)
    assert np.array_equal(uo, go) and np.array_equal(ui, gi)
    Fg, Kg = gpu.assemble(x, elastic_folds=True, exact_hinge=exact)
    assert np.abs(Fg - Fr).max() <= 1e-10 * np.abs(Fr).max()
    assert np.abs(Kg - Kr).max() <= 1e-12 * np.abs(Kr).max()
    # with every fold elastic the residual differs by fold 0's hinge force
    gpu.set_fold_rigid(0, False)
    gpu.init_assembly()
    F_all, _ = gpu.assemble(x, elastic_folds=True, tangent=False)
    gpu.set_fold_rigid(0, True)
    gpu.init_assembly()
    assert np.abs(F_all - Fg).max() > 0.0


def test_constraint_rows_rigid_fold_only(fv):
    sh, orc, gpu = fv
    x = W.displacement(sh.nn, 8, 1e-3)
    rows, cols = gpu.constraint_layout(rigid=False)
    c_o, r_o, col_o, v_o = orc.constraints(x, flag_rig=False)
    n = [f.left.n for f in sh.folds]
    assert c_o.size == 4 * n[0] + 3 * n[1] + 3 * n[2]      # angle rows for fold 0 only
    assert np.array_equal(rows, r_o) and np.array_equal(cols, col_o)
    c_g, v_g = gpu.constraints(x)
    ang = 4 * np.arange(n[0]) + 3                              # fold 0: 3 tie rows, then its angle row
    lin = np.setdiff1d(np.arange(c_o.size), ang)
    # tie rows are linear in Psi (1e-12); angle rows go through F0, i.e. grad Psi
    # of the fold GPs next to the crossing and the arc, known to ~1e-11 (1e-10 rad)
    assert np.abs(c_g[lin] - c_o[lin]).max() <= 1e-12
    assert np.abs(c_g[ang] - c_o[ang]).max() <= 1e-10
    assert np.abs(v_g - v_o).max() <= 1e-10 * max(1.0, np.abs(v_o).max())
