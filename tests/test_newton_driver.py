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
