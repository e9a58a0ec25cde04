"""The host Newton driver (continuation.py) and the oracle's restated LineSearch
(oracle/newton.py) on the reference's own solver-layer known answers
(tests/consumer/main.cpp:27-118, :129-206): a constrained cubic spring with a
closed-form equilibrium and a follower torque with a closed-form path.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import continuation as CS
import newton as ON


class CubicSpring:
    """Two DOFs: x0 on a grounded spring k0, joined to x1 by f(d) = kc d + a d^3,
    d = x1 - x0; load lambda on x1; constraint x0 = 0. At lambda = 2 with
    kc = a = 1 the equilibrium is x1 = 1 (tests/consumer/main.cpp:1-60)."""
    k0, kc, a = 10.0, 1.0, 1.0
    num_dofs = 2

    def internal_force(self, x):
        d = x[1] - x[0]
        f = self.kc * d + self.a * d ** 3
        return np.array([self.k0 * x[0] - f, f])

    def assemble(self, x):
        d = x[1] - x[0]
        kt = self.kc + 3 * self.a * d * d
        K = sp.csc_matrix(np.array([[self.k0 + kt, -kt], [0.0, kt]]))   # upper triangle
        return K, self.internal_force(x)

    def external_force(self, x):
        return np.array([0.0, 1.0])

    def load_stiffness(self, x):
        return None

    def constraints(self, x):
        return np.array([x[0]]), sp.csr_matrix(np.array([[1.0, 0.0]]))


class FollowerSpring:
    """One rotation theta: spring k theta under lambda m sin(beta - theta); exact
    path lambda(theta) = k theta / (m sin(beta - theta)) (tests/consumer/main.cpp:62-118)."""
    k, m, beta = 2.0, 1.0, 0.5 * math.pi
    num_dofs = 1

    def __init__(self, use_load_stiffness):
        self.use = use_load_stiffness
        self.assemblies = 0

    def internal_force(self, x):
        return np.array([self.k * x[0]])

    def assemble(self, x):
        self.assemblies += 1
        return sp.csc_matrix(np.array([[self.k]])), self.internal_force(x)

    def external_force(self, x):
        return np.array([self.m * math.sin(self.beta - x[0])])

    def load_stiffness(self, x):
        if not self.use:
            return None
        return sp.csc_matrix(np.array([[-self.m * math.cos(self.beta - x[0])]]))

    def constraints(self, x):
        return np.zeros(0), sp.csr_matrix((0, 1))

    def path_error(self, lam, th):
        return abs(lam - self.k * th / (self.m * math.sin(self.beta - th)))


def test_linesearch_cubic_spring():
    x = np.zeros(2)
    rep = CS.ContinuationSolver().line_search(CubicSpring(), 2.0, x)
    assert rep.converged
    assert abs(x[1] - 1.0) <= 1e-8
    assert abs(x[0]) <= 1e-12


def test_oracle_linesearch_cubic_spring():
    x = np.zeros(2)
    ok, _ = ON.line_search(CubicSpring(), 2.0, x)
    assert ok
    assert abs(x[1] - 1.0) <= 1e-8
    assert abs(x[0]) <= 1e-12


def test_linesearch_follower_torque():
    its, th = {}, {}
    for use in (True, False):
        prob = FollowerSpring(use)
        x = np.zeros(1)
        rep = CS.ContinuationSolver().line_search(prob, 1.0, x)
        assert rep.converged
        assert prob.path_error(1.0, x[0]) <= 1e-7
        its[use], th[use] = prob.assemblies, x[0]
    assert its[True] < its[False]
    assert abs(th[True] - th[False]) <= 1e-8


def test_load_path_cubic_spring_increments():
    """Proportional increments to lambda = 2: every increment converges and each
    state lies on x1 + x1^3 = lambda."""
    prob = CubicSpring()
    res = CS.load_path(prob, increments=8, lam_total=2.0)
    assert res.all_converged and len(res.converged) == 8
    assert abs(res.x[1] + res.x[1] ** 3 - 2.0) <= 1e-8


def test_rollback_on_divergence():
    """A problem whose residual is beyond div rolls back to the start state
    (ContinuationSolver.h:235-239)."""
    prob = CubicSpring()
    x = np.array([0.0, 0.5])
    rep = CS.ContinuationSolver().line_search(prob, 1e9, x)
    assert not rep.converged
    assert np.array_equal(x, [0.0, 0.5])


def test_symmetric_completion_ignores_stored_lower_entries():
    """The assembled K stores full 3 x 3 diagonal node blocks; the solver must read
    only the upper triangle (selfadjointView<Upper>, MUMPS sym = 2)."""
    K = sp.csc_matrix(np.array([[4.0, 1.0], [7.0, 3.0]]))     # 7.0 is a stored lower entry
    F = CS.full_from_upper(K).toarray()
    assert np.array_equal(F, [[4.0, 1.0], [1.0, 3.0]])
    assert np.array_equal(ON._sym(K).toarray(), [[4.0, 1.0], [1.0, 3.0]])


def test_independent_rows_drop_duplicates():
    """InitializeTopology keeps an independent row set (column-pivoted QR of J^T)."""
    J = sp.csr_matrix(np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 0.0, 2.0]]))
    s = CS.ContinuationSolver()
    s.initialize_topology(J)
    assert s.ind_rows.size == 3
    assert np.linalg.matrix_rank(J.toarray()[s.ind_rows]) == 3
    assert ON._independent_rows(J).size == 3


# ---- ArcLength (ContinuationSolver.h:329-538) -------------------------------------
def _arclength_cubic(driver):
    acc = []
    if driver == "host":
        s = CS.ContinuationSolver()
        s.psi, s.ds, s.dsmax = 1.0, 0.2, 0.5
        s.arc_length(CubicSpring(), 0.0, np.zeros(2), 5.0, lambda k, lam, x, W: acc.append((lam, x.copy())))
    else:
        ON.arc_length(CubicSpring(), 0.0, np.zeros(2), 5.0, lambda k, lam, x: acc.append((lam, x.copy())),
                      ds=0.2, dsmax=0.5, psi=1.0)
    return acc


@pytest.mark.parametrize("driver", ["host", "oracle"])
def test_arclength_cubic_spring(driver):
    """tests/consumer/main.cpp:138-160: psi = 1, ds = 0.2, dsmax = 0.5 to lambda = 5;
    every accepted step on x1 + x1^3 = lambda (1e-8) with x0 = 0 (1e-12), and the
    last one past lambda = 5."""
    acc = _arclength_cubic(driver)
    assert len(acc) > 0 and acc[-1][0] > 5.0
    assert max(abs(x[1] + x[1] ** 3 - lam) for lam, x in acc) <= 1e-8
    assert max(abs(x[0]) for _, x in acc) <= 1e-12


def test_arclength_drivers_agree():
    """The host driver and the oracle's restatement take the same steps."""
    a, b = _arclength_cubic("host"), _arclength_cubic("oracle")
    assert len(a) == len(b)
    for (la, xa), (lb, xb) in zip(a, b):
        assert abs(la - lb) <= 1e-12 * max(1.0, abs(lb)) and np.abs(xa - xb).max() <= 1e-12


def test_arclength_follower_torque():
    """tests/consumer/main.cpp:189-216: psi = 1, ds = dsmax = 0.2 to lambda = 2, on
    the exact path lambda = k theta / (m sin(beta - theta)) within 1e-7 with and
    without the load stiffness, fewer tangent assemblies per step with it."""
    per_step = {}
    for use in (True, False):
        prob = FollowerSpring(use)
        s = CS.ContinuationSolver()
        s.psi, s.ds, s.dsmax = 1.0, 0.2, 0.2
        acc = []
        s.arc_length(prob, 0.0, np.zeros(1), 2.0, lambda k, lam, x, W: acc.append((lam, x[0])))
        assert len(acc) > 0 and acc[-1][0] > 2.0
        assert max(prob.path_error(lam, th) for lam, th in acc) <= 1e-7
        per_step[use] = prob.assemblies / len(acc)
    assert per_step[True] < per_step[False]


class SnapSpring:
    """One DOF with a non-monotone spring, F_int = x^3 / 3 - x^2 + 0.8 x (stiffness
    (x - 1)^2 - 0.2), under a dead load lambda: the equilibrium path
    lambda(x) = F_int(x) rises to a limit point at x = 1 - sqrt(0.2), falls to a
    second one at 1 + sqrt(0.2) and rises again. Arc-length follows it through both, reversing the load increment after
    each limit point."""
    num_dofs = 1

    def internal_force(self, x):
        return np.array([x[0] ** 3 / 3.0 - x[0] ** 2 + 0.8 * x[0]])

    def assemble(self, x):
        return sp.csc_matrix(np.array([[x[0] ** 2 - 2.0 * x[0] + 0.8]])), self.internal_force(x)

    def external_force(self, x):
        return np.array([1.0])

    def load_stiffness(self, x):
        return None

    def constraints(self, x):
        return np.zeros(0), sp.csr_matrix((0, 1))


def test_arclength_through_limit_points():
    s = CS.ContinuationSolver()
    s.psi, s.ds, s.dsmax = 1.0, 0.05, 0.1
    acc = []
    rep = s.arc_length(SnapSpring(), 0.0, np.zeros(1), 1.0, lambda k, lam, x, W: acc.append((lam, x[0])))
    lams = np.array([a for a, _ in acc]); xs = np.array([b for _, b in acc])
    assert rep.stop == "lambda_tot" and lams[-1] > 1.0
    assert np.abs(xs ** 3 / 3 - xs ** 2 + 0.8 * xs - lams).max() <= 1e-8
    assert np.all(np.diff(xs) > 0)                       # the path never turns back in x
    assert np.any(np.diff(lams) < 0) and rep.reversals >= 1
    x1, x2 = 1 - math.sqrt(0.2), 1 + math.sqrt(0.2)
    assert xs.min() < x1 and xs.max() > x2


def test_arclength_restart_resumes_the_run(tmp_path):
    """RestartManager: a run stopped after 5 steps and resumed from its .res record
    ends in the uninterrupted run's state, with the same work. To round-off: the
    KKT factorization's scaling and matching come from the values of its first
    analysis, which the resumed run performs at the restart state."""
    full = CS.ContinuationSolver()
    full.psi, full.ds, full.dsmax = 1.0, 0.2, 0.5
    ref = full.arc_length(CubicSpring(), 0.0, np.zeros(2), 5.0)
    path = str(tmp_path / "cubic.res")
    first = CS.ContinuationSolver()
    first.psi, first.ds, first.dsmax, first.kmax, first.restart_file = 1.0, 0.2, 0.5, 5, path
    rep1 = first.arc_length(CubicSpring(), 0.0, np.zeros(2), 5.0)
    assert rep1.stop == "kmax" and len(rep1.steps) == 5
    second = CS.ContinuationSolver()
    second.psi, second.ds, second.dsmax, second.restart_file = 1.0, 0.2, 0.5, path
    rep2 = second.arc_length(CubicSpring(), 0.0, np.zeros(2), 5.0)
    assert rep2.restarted_from == 6
    assert len(rep1.steps) + len(rep2.steps) == len(ref.steps)
    assert rep2.lam == pytest.approx(ref.lam, rel=1e-13)
    assert np.abs(rep2.x - ref.x).max() <= 1e-13 * np.abs(ref.x).max()
    assert rep2.steps[-1].W_ext == pytest.approx(ref.steps[-1].W_ext, rel=1e-12)
