"""Pins of the oracle's constraint rows and external load against closed forms.

Constraint rows restate Constraints::evaluateDirichlet / evaluateFold
(plate/Constraints.h:24-52, 247-308); the load restates
Meshless::computeExternalForce (plate/Meshless.h:60-100). Quadratic RKPM
reproduces polynomials of degree <= 2 exactly, which fixes the values below.
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
import workloads as W
from _cases import RHO_FACTOR


@pytest.fixture(scope="module")
def loaded():
    sh = W.with_edge_moment_load(W.config1(rho_factor=RHO_FACTOR))
    return sh, O.plate_from_sheet(sh, nthreads=4)


def _field(sh, fu, fv, fw):
    x = np.empty(3 * sh.nn)
    x[0::3] = fu(sh.X, sh.Y); x[1::3] = fv(sh.X, sh.Y); x[2::3] = fw(sh.X, sh.Y)
    return x


def test_row_layout(loaded):
    """6 Dirichlet rows (u,v,w at (0,0); v,w at (1,0); w at (0,1)) then 3 tie
    rows per crease GP; 4 per GP with rigid folds."""
    sh, p = loaded
    x = np.zeros(3 * sh.nn)
    c, r, col, v = p.constraints(x, flag_rig=False)
    ngp = sh.folds[0].left.n
    assert c.size == 6 + 3 * ngp
    assert np.all(np.diff(r) >= 0)          # rows are pushed in order
    c2, r2, _, _ = p.constraints(x, flag_rig=True)
    assert c2.size == 6 + 4 * ngp
    assert np.abs(c).max() == 0.0


def test_dirichlet_rows_on_linear_field(loaded):
    sh, p = loaded
    a = np.array([0.3, -0.2, 0.7]); B = np.array([[0.1, -0.4], [0.25, 0.05], [-0.6, 0.35]])
    x = _field(sh, *[lambda X, Y, k=k: a[k] + B[k, 0] * X + B[k, 1] * Y for k in range(3)])
    c, r, col, v = p.constraints(x)
    u = lambda px, py: a + B @ np.array([px, py])
    expect = [u(0, 0)[0], u(0, 0)[1], u(0, 0)[2], u(1, 0)[1], u(1, 0)[2], u(0, 1)[2]]
    assert np.abs(c[:6] - expect).max() <= 1e-13
    # each Dirichlet row holds Psi of its point: partition of unity
    for row in range(6):
        assert abs(v[r == row].sum() - 1.0) <= 1e-13
    # columns carry the constrained component
    comps = [0, 1, 2, 1, 2, 2]
    for row, comp in enumerate(comps):
        assert np.all(col[r == row] % 3 == comp)


def test_tie_rows_measure_the_offset_jump(loaded):
    """For a linear field u = a + B X, u_L - u_R = B (X_L - X_R) = 2e-7 B n."""
    sh, p = loaded
    a = np.array([0.3, -0.2, 0.7]); B = np.array([[0.1, -0.4], [0.25, 0.05], [-0.6, 0.35]])
    x = _field(sh, *[lambda X, Y, k=k: a[k] + B[k, 0] * X + B[k, 1] * Y for k in range(3)])
    c, r, col, v = p.constraints(x)
    fl = sh.folds[0]
    d = np.stack([fl.left.gx - fl.right.gx, fl.left.gy - fl.right.gy])   # 2e-7 n
    jump = (B @ d).T.reshape(-1)
    assert np.abs(c[6:] - jump).max() <= 1e-12     # reproduction holds to ~1e-13 of |u|
    # tie row entries: +Psi_L and -Psi_R, each summing to one
    for row in range(6, 9):
        vv = v[r == row]
        assert abs(vv[vv > 0].sum() - 1.0) <= 1e-12 or abs(vv.sum()) <= 1e-12


def test_rigid_angle_row(loaded):
    """Rigid fold of the + side by phi: the angle row reads phi - theta0 and its
    Jacobian row matches central differences of c."""
    sh, p = loaded
    phi = 0.6
    x = W.rigid_fold_displacement(sh, phi, seed=5, sigma=0.0)
    c, r, col, v = p.constraints(x, flag_rig=True)
    ngp = sh.folds[0].left.n
    ang = c[6 + 3 + np.arange(ngp) * 4]
    assert np.abs(np.abs(ang) - phi).max() <= 1e-9
    # Jacobian of one angle row by central differences
    row = 6 + 3
    rng = np.random.default_rng(2)
    dv = rng.standard_normal(x.size); dv /= np.linalg.norm(dv)
    h = 1e-6
    cp, *_ = p.constraints(x + h * dv, flag_rig=True)
    cm, *_ = p.constraints(x - h * dv, flag_rig=True)
    fd = (cp[row] - cm[row]) / (2 * h)
    an = np.sum(v[r == row] * dv[col[r == row]])
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))


def test_edge_moment_work_closed_form(loaded):
    """F_ext . w for w = x^2/2 is sum over edge GPs of W m (dw/dn_out) = m (only
    the x = 1 edge has a slope, outward normal +x, length 1); w = y^2/2 and w = x
    do no work (slopes cancel or vanish)."""
    sh, p = loaded
    m = sh.boundaries[3].moment
    F = p.external_force()
    assert np.all(F[0::3] == 0.0) and np.all(F[1::3] == 0.0)
    Fw = F[2::3]
    assert abs(Fw @ (0.5 * sh.X ** 2) - m) <= 1e-12 * m
    assert abs(Fw @ (0.5 * sh.Y ** 2)) <= 1e-12 * m
    assert abs(Fw @ sh.X) <= 1e-12 * m


def test_dead_force_partition_of_unity():
    """A dead force f on a domain adds f times the domain weight in total."""
    sh = W.with_edge_moment_load(W.config1(rho_factor=RHO_FACTOR))
    b = list(sh.boundaries)
    b[1] = dataclasses.replace(b[1], force=(0.0, 0.0, -2.5))        # point (1, 0)
    b[3] = dataclasses.replace(b[3], force=(0.1, 0.0, 0.0), moment=0.0)   # edge x = 0
    sh = dataclasses.replace(sh, boundaries=b)
    p = O.plate_from_sheet(sh, nthreads=4)
    F = p.external_force().reshape(-1, 3)
    assert abs(F[:, 0].sum() - 0.1 * 1.0) <= 1e-13
    assert abs(F[:, 2].sum() - (-2.5)) <= 1e-12 + 1e-12 * abs(sh.boundaries[4].moment)
