"""Progressive creasing on the CPU: the oracle's run of the reference's smoke
example against the answer its README states, the stepping helpers, and the
oracle's energies and crease angles pinned by closed forms and by the
residual (F_int = dE/dx, checked by central differences).

The smoke answer (examples/smoke_centre_crease/README.md:14-20) is quoted to
4 significant digits "to round-off"; the bar is half a unit of the last
printed digit. The same run's per-step states are the golden fixture
tests/golden/progressive_smoke.npz (make_progressive_fixture.py) for the
device run.
"""
import math
import os

import numpy as np
import pytest

import oracle as O
import progressive as PR
import workloads as W
from _cases import RHO_FACTOR
from _progressive_oracle import SMOKE_KNOWN, SMOKE_MAX_ITERS, run_oracle_smoke

INPUTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_inputs")
FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "progressive_smoke.npz")


def within_printed_digits(value, quoted, digits=4):
    """|value - quoted| <= half a unit in the last of `digits` significant digits."""
    unit = 10.0 ** (math.floor(math.log10(abs(quoted))) - (digits - 1))
    return abs(value - quoted) <= 0.5 * unit


def test_oracle_smoke_reaches_the_known_answer():
    recs, x = run_oracle_smoke(INPUTS)
    assert [(r.a, r.b) for r in recs] == [(5, 5), (4, 6), (3, 7), (2, 8), (1, 9), (0, 10)]
    assert all(r.iters <= SMOKE_MAX_ITERS and r.substeps == 0 for r in recs)
    last = recs[-1]
    assert last.step == SMOKE_KNOWN["step"]
    for k in ("E_mem", "E_ben", "E_hin"):
        assert within_printed_digits(getattr(last, k), SMOKE_KNOWN[k]), (k, getattr(last, k))
    assert within_printed_digits(last.theta_mean_deg, SMOKE_KNOWN["theta_mean_deg"])
    ref = np.load(FIXTURE)
    # the fixture is this run (to Newton round-off: the threaded oracle sums
    # in a different order from run to run)
    assert np.allclose([r.E_ben for r in recs], ref["E_ben"], rtol=1e-9, atol=0)
    assert np.abs(x - ref["x"]).max() <= 1e-9 * np.abs(ref["x"]).max()


def test_schedules():
    s = lambda mode, step, N, skip=(): [(t.a, t.b) for t in PR.schedule(mode, step, 4, N, skip)]
    assert s("middle_out", 1, 10) == [(5, 5), (4, 6), (3, 7), (2, 8), (1, 9), (0, 10)]
    assert s("middle_out", 2, 6) == [(3, 3), (1, 5), (0, 6)]
    assert s("one_end", 2, 5) == [(0, 0), (0, 2), (0, 4), (0, 5)]
    assert s("one_end", 1, 4, skip=(2,)) == [(0, 0), (0, 1), (0, 3), (0, 4)]
    assert s("one_end_reverse", 2, 5) == [(5, 5), (3, 5), (1, 5), (0, 5)]
    ei = PR.schedule("ends_in", 1, 4, 4)
    assert [(t.a, t.b, t.outside) for t in ei] == [(0, 4, True), (1, 3, True), (2, 2, True)]
    aa = PR.schedule("all_at_once", 1, 4, 7)
    assert [(t.a, t.b) for t in aa] == [(0, 7)] * 5 and [t.frac for t in aa] == [0, 0.25, 0.5, 0.75, 1.0]
    with pytest.raises(ValueError):
        PR.schedule("sideways", 1, 4, 4)


def test_truncate_and_arc_length():
    full = W.Crease(np.array([0.0, 1.0, 1.0, 2.0]), np.array([0.0, 0.0, 1.0, 1.0]), True, False)
    S = PR.cumulative_arc_length(full)
    assert S.tolist() == [0.0, 1.0, 2.0, 3.0]
    assert PR.project_arc_length(1.2, 0.5, full, S) == pytest.approx(1.5)
    assert PR.project_arc_length(-1.0, 0.0, full, S) == 0.0
    c = PR.truncate(full, 0, 2)
    assert c.vx.tolist() == [0.0, 1.0, 1.0] and c.ext_s and not c.ext_e
    c = PR.truncate(full, 1, 3)
    assert not c.ext_s and not c.ext_e


def test_fold_angle_statistics_across_the_branch_cut():
    """Angles straddling +-180 deg average near 180, not near 0
    (plate/FoldAngleUtil.h:5-24)."""
    th = np.radians([179.0, -179.0, 178.0])
    m, mx = PR.active_fold_angles(th, np.ones(3), np.ones(3))
    assert abs(m - (179.0 + 181.0 + 178.0) / 3.0) < 1e-9 and mx == pytest.approx(181.0)
    m, mx = PR.active_fold_angles(np.radians([10.0, 20.0, 99.0]), np.array([1.0, 3.0, 1.0]), [1, 1, 0])
    assert m == pytest.approx(17.5) and mx == pytest.approx(20.0)
    assert PR.active_fold_angles(np.zeros(3), np.ones(3), np.zeros(3)) == (0.0, 0.0)


@pytest.fixture(scope="module")
def c1():
    sh = W.config1(rho_factor=RHO_FACTOR)
    return sh, O.plate_from_sheet(sh, nthreads=4)


def test_energy_gradient_is_the_residual(c1):
    """F_int (pinned against the reference elsewhere) is dE/dx of
    E_mem + E_ben + E_hin: central differences along random directions."""
    sh, orc = c1
    x = W.displacement(sh.nn, 7, 2e-3)
    F = orc.internal_force(x, flag_rig=False)
    F_face = orc.internal_force(x, flag_rig=True)          # no hinge terms
    h = 1e-7                                                # truncation ~ h^2: 3e-6 at 1e-5, 4e-10 here
    for seed in (1, 2):
        v = W.displacement(sh.nn, 100 + seed, 1.0)
        fd = (np.array(orc.energies(x + h * v)) - np.array(orc.energies(x - h * v))) / (2 * h)
        # bars on the scale of the directional derivative, |F| |v| (F.v itself
        # can cancel)
        assert abs(fd.sum() - F @ v) <= 1e-8 * np.linalg.norm(F) * np.linalg.norm(v), (fd, F @ v)
        Fh = F - F_face
        assert abs(fd[2] - Fh @ v) <= 1e-6 * np.linalg.norm(Fh) * np.linalg.norm(v), (fd[2], Fh @ v)


def test_energies_vanish_under_rigid_motion(c1):
    sh, orc = c1
    a = 0.3
    X, Y = sh.X, sh.Y
    x = np.zeros(3 * sh.nn)
    x[0::3] = X * math.cos(a) - Y * math.sin(a) - X + 0.2
    x[1::3] = X * math.sin(a) + Y * math.cos(a) - Y - 0.1
    x[2::3] = 0.05
    Em, Eb, Eh = orc.energies(x)
    scale = sum(orc.energies(W.displacement(sh.nn, 3, 1e-3)))
    assert max(abs(Em), abs(Eb), abs(Eh)) <= 1e-9 * scale


def test_rigid_fold_hinge_energy_and_angles(c1):
    """A rigid fold by alpha about the straight crease: theta = -alpha at every
    crease GP (the + side turned by alpha), E_mem = E_ben = 0, and
    E_hin = 1/2 Kt (alpha - Ang0)^2 x crease length."""
    sh, orc = c1
    alpha = 0.4
    x = W.rigid_fold_displacement(sh, alpha, seed=0, sigma=0.0)
    th = orc.fold_angles(0, x)
    assert np.abs(np.abs(th) - alpha).max() < 1e-9
    f = sh.folds[0]
    Em, Eb, Eh = orc.energies(x)
    L = float(np.sum(f.left.w))
    expect = 0.5 * f.Kt * (th[0] - f.Ang0) ** 2 * L
    assert Eh == pytest.approx(expect, rel=1e-9)
    assert max(abs(Em), abs(Eb)) <= 1e-9 * Eh
    orc.set_fold_gp_state(0, np.full(th.size, th[0]), np.ones(th.size))
    assert orc.energies(x)[2] < 1e-18 * max(1.0, Eh)
    act = np.zeros(th.size); act[::2] = 1
    orc.set_fold_gp_state(0, None, act)
    assert orc.energies(x)[2] == pytest.approx(0.5 * f.Kt * th[0] ** 2 * float(np.sum(f.left.w[::2])), rel=1e-9)
    orc.set_fold_gp_state(0, None, None)
