"""Input generators: sizes, quadrature weights, RNG, crease geometry."""
import math

import numpy as np

import workloads as W


def test_splitmix_and_normals_are_fixed():
    z = W.splitmix64(1, 3)
    # SplitMix64 reference outputs for state 1 (first three draws)
    assert [int(v) for v in z] == [0x910A2DEC89025CC1, 0xBEEB8DA1658EEC67, 0xF893A2EEFB32555E]
    u = W.uniform01(1, 1000)
    assert u.min() > 0.0 and u.max() <= 1.0 and abs(u.mean() - 0.5) < 0.03
    x = W.normal(7, 20000, 2.0)
    assert abs(x.std() - 2.0) < 0.05 and abs(x.mean()) < 0.05
    assert np.array_equal(W.normal(7, 11), W.normal(7, 12)[:11])


def test_config1_sizes_and_weights():
    sh = W.config1(rho_factor=3.75)
    assert sh.nn == 441 and sh.face.n == 6400 and sh.n_split == 0
    assert abs(sh.face.w.sum() - 1.0) < 1e-13
    assert abs(sh.rho - 3.75 * 0.05) < 1e-15
    f = sh.folds[0]
    assert f.left.n == 20 and abs(f.Kt - 1e-8) < 1e-22
    assert np.allclose(f.left.gx, 0.5 - 1e-7) and np.allclose(f.right.gx, 0.5 + 1e-7)
    assert np.allclose(f.left.tx, 0) and np.allclose(f.left.ty, 1) and np.allclose(f.left.nx, -1)
    assert abs(f.left.w.sum() - 1.0) < 1e-13


def test_config2_arc():
    sh = W.config2(rho_factor=3.75, n_side=41, nseg=64)
    cr = sh.creases[0]
    r = np.hypot(cr.vx - 0.5, cr.vy + 0.1)
    assert np.allclose(r, 0.4) and abs(cr.vy[0]) < 1e-15 and abs(cr.vy[-1]) < 1e-15
    assert cr.ext_s and cr.ext_e
    L = np.hypot(np.diff(cr.vx), np.diff(cr.vy)).sum()
    assert abs(sh.folds[0].left.w.sum() - L) < 1e-13
    assert abs(sh.face.w.sum() - 1.0) < 1e-12


def test_split_on_crease():
    """A face GP exactly on a crease becomes two half-weight GPs at +-1e-7."""
    face = W.GPSet(np.array([0.5, 0.3]), np.array([0.25, 0.25]), np.array([1.0, 1.0]))
    cr = W.Crease(np.array([0.5, 0.5]), np.array([0.0, 1.0]))
    out, n = W.split_on_creases(face, [cr])
    assert n == 1 and out.n == 3
    assert np.allclose(out.w, [0.5, 1.0, 0.5])
    assert np.allclose(out.gx, [0.5 - 1e-7, 0.3, 0.5 + 1e-7])


def test_rigid_fold_keeps_plus_side_distances():
    sh = W.config1(rho_factor=3.75)
    x = W.rigid_fold_displacement(sh, math.radians(60), seed=0, sigma=0.0)
    P = np.stack([sh.X + x[0::3], sh.Y + x[1::3], x[2::3]], 1)
    plus = sh.X <= 0.5 + 1e-12
    i, j = np.nonzero(plus)[0][[3, 200]]
    d0 = math.hypot(sh.X[i] - sh.X[j], sh.Y[i] - sh.Y[j])
    assert abs(np.linalg.norm(P[i] - P[j]) - d0) < 1e-12
    assert np.all(x[0::3][~plus] == 0)
