"""The host KKT solver (libcurvami_kkt, the MUMPS replacement) on the CPU:

- solutions of [K J^T; J 0] equal scipy's SuperLU solve of the assembled
  symmetric matrix to the conditioning of the system (100 eps cond), with a
  normwise backward error below 1e-13;
- the reported negative pivot count equals the number of negative eigenvalues
  of the KKT matrix (dense eigvalsh), stable (K positive definite on ker J:
  exactly the constraint count) and unstable states;
- the symbolic analysis runs once per pattern and again only when K's or J's
  pattern changes (MumpsSymmetricSolver.h:184-196);
- J's duplicate triplets are summed, K's stored entries below the diagonal
  (the lower halves of the diagonal node blocks) are ignored;
- bad input is refused with a status and message, before any factorization;
- solutions are bit-identical run to run and across thread counts.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import curvami_gpu as G
import oracle as O
import workloads as W


@pytest.fixture(scope="module")
def system():
    sh = W.with_edge_moment_load(W.config1(rho_factor=3.75))
    orc = O.plate_from_sheet(sh, nthreads=4)
    out = {}
    for name, sigma in (("noise", 1e-3), ("rest", 0.0)):
        x = W.displacement(sh.nn, 1, sigma)
        outer, inner, K, _ = orc.assemble(x, flag_rig=False, nthreads=4)
        _, r, col, v = orc.constraints(x, flag_rig=False)
        out[name] = (outer, inner, K, r, col, v)
    return out


def _dense_kkt(outer, inner, K, r, col, v):
    n = outer.size - 1
    m = int(r.max()) + 1
    U = sp.csc_matrix((K, inner, outer), shape=(n, n))
    Ks = sp.triu(U) + sp.triu(U, 1).T
    J = sp.coo_matrix((v, (r, col)), shape=(m, n)).tocsr()
    return sp.bmat([[Ks, J.T], [J, None]], format="csc"), n, m


@pytest.mark.parametrize("state", ["noise", "rest"])
def test_solution_and_inertia(system, state):
    outer, inner, K, r, col, v = system[state]
    A, n, m = _dense_kkt(outer, inner, K, r, col, v)
    rhs = np.random.default_rng(0).standard_normal((n + m, 2))
    ref = spla.splu(A).solve(rhs)
    s = G.KKT(4)
    s.set_pattern(outer, inner, m, r, col)
    neg, pos = s.factorize(K, v)
    sol = s.solve(rhs)
    ev = np.linalg.eigvalsh(A.toarray())
    cond = np.abs(ev).max() / np.abs(ev).min()
    err = np.abs(sol - ref).max() / np.abs(ref).max()
    # normwise backward error |A x - b| / (|A| |x| + |b|) (inf-norms)
    anorm = abs(A).sum(axis=1).max()
    res = np.abs(A @ sol - rhs).max() / (anorm * np.abs(sol).max() + np.abs(rhs).max())
    print(f"{state}: neg {neg} (dense {int((ev < 0).sum())}), cond {cond:.2e}, vs SuperLU {err:.2e}, "
          f"backward error {res:.2e}")
    assert neg == int((ev < 0).sum()) and pos == int((ev > 0).sum())
    if state == "rest":
        assert neg == m          # stable: the reduced Hessian is positive definite
    assert err <= 1e2 * np.finfo(float).eps * cond and res <= 1e-13


def test_analysis_reused_until_pattern_changes(system):
    outer, inner, K, r, col, v = system["noise"]
    m = int(r.max()) + 1
    s = G.KKT(2)
    for scale in (1.0, 1.5, 2.0):
        s.set_pattern(outer, inner, m, r, col)
        s.factorize(K * scale, v)
    st = s.stats()
    assert st["analyses"] == 1 and st["factorizations"] == 3
    keep = np.arange(r.size) % 7 != 0          # J's pattern changes
    s.set_pattern(outer, inner, m, r[keep], col[keep])
    s.factorize(K, v[keep])
    assert s.stats()["analyses"] == 2


def test_duplicates_summed_and_lower_entries_ignored(system):
    outer, inner, K, r, col, v = system["noise"]
    A, n, m = _dense_kkt(outer, inner, K, r, col, v)
    rhs = np.random.default_rng(1).standard_normal(n + m)
    s = G.KKT(2)
    s.set_pattern(outer, inner, m, r, col)
    s.factorize(K, v)
    x1 = s.solve(rhs)
    # J's triplets split in halves (duplicates), K's strictly lower stored entries scrambled
    r2, c2, v2 = np.concatenate([r, r]), np.concatenate([col, col]), np.concatenate([0.5 * v, 0.5 * v])
    colidx = np.repeat(np.arange(n), np.diff(outer))
    Kx = np.where(inner > colidx, 1e6, K)
    t = G.KKT(2)
    t.set_pattern(outer, inner, m, r2, c2)
    t.factorize(Kx, v2)
    x2 = t.solve(rhs)
    assert np.abs(x2 - x1).max() <= 1e-9 * np.abs(x1).max()


def test_bad_input_refused(system):
    outer, inner, K, r, col, v = system["noise"]
    m = int(r.max()) + 1
    s = G.KKT(1)
    with pytest.raises(G.CvmError) as e:
        s.factorize(K, v)
    assert e.value.code == 7
    with pytest.raises(G.CvmError) as e:
        s.set_pattern(outer, inner, m, r, col + 10 ** 6)
    assert e.value.code == 1 and "outside" in str(e.value)
    with pytest.raises(G.CvmError) as e:
        s.solve(np.zeros(outer.size - 1 + m))
    assert e.value.code == 7
    # column pointers must start at zero
    with pytest.raises(G.CvmError) as e:
        s.set_pattern(outer - 1, inner, m, r, col)
    assert e.value.code == 1 and "bad pattern" in str(e.value)


def test_bit_reproducible_across_threads(system):
    outer, inner, K, r, col, v = system["noise"]
    m = int(r.max()) + 1
    rhs = np.random.default_rng(2).standard_normal(outer.size - 1 + m)
    sols = []
    for T in (4, 4, 1, 2):
        s = G.KKT(T)
        s.set_pattern(outer, inner, m, r, col)
        s.factorize(K, v)
        sols.append(s.solve(rhs))
    for b in sols[1:]:
        assert np.array_equal(b.view(np.int64), sols[0].view(np.int64))


def test_nearly_singular_saddle_point_falls_back_to_lu():
    """The reference's ls_miura configuration at its flat start (four rigid folds
    and two cuts held by 190 constraint rows, cubic basis, rho = 6h): the
    symmetric factorization's probe solve misses its backward-error bar, the LU
    takes over, and the solve agrees with SuperLU's (inertia reported unknown)."""
    import json
    import os
    import continuation as CS
    import mesh_setup as MS
    import newton as N
    inputs = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_inputs")
    cfg = json.load(open(os.path.join(inputs, "ls_miura.json")))
    ms = MS.sheet_from_config(cfg, base_dir=inputs, mesh_file=os.path.join(inputs, "Miura.msh"))
    prob = N.OracleProblem(O.plate_from_sheet(ms.sheet, nthreads=4), flag_rig=True, nthreads=4)
    x = np.zeros(prob.num_dofs)
    c, J = prob.constraints(x)
    s = CS.ContinuationSolver(kkt=CS.SuperLUKKTSolver())
    s.initialize_topology(J)
    ci, Ji = s.filter_rows(c, J)
    K, _ = prob.assemble(x)
    n, m = prob.num_dofs, ci.size
    rhs = np.zeros(n + m)
    rhs[:n] = -CS.sym_matvec(K, CS.ConstraintProjector(Ji).corrector(ci, n))
    ref = CS.SuperLUKKTSolver().solve(K, Ji, rhs)
    kk = CS.KKTSolver(4)
    sol = kk.solve(K, Ji, rhs)
    st = kk.stats()
    print(st)
    assert st["lu_fallbacks"] == 1 and kk.neg_pivots == -1
    # the displacement part agrees; the multipliers of this ill-conditioned
    # system are determined less sharply, so both solutions are held to a
    # backward-error bar instead
    assert np.abs(sol[:n] - ref[:n]).max() <= 1e-9 * np.abs(ref[:n]).max()
    A = sp.bmat([[CS.full_from_upper(K), Ji.T], [Ji, None]], format="csc")
    anorm = abs(A).sum(axis=1).max()
    for y in (sol, ref):
        assert np.abs(A @ y - rhs).max() <= 1e-14 * (anorm * np.abs(y).max() + np.abs(rhs).max())


def test_well_conditioned_systems_keep_ldlt(system):
    outer, inner, K, r, col, v = system["noise"]
    m = int(r.max()) + 1
    s = G.KKT(2)
    s.set_pattern(outer, inner, m, r, col)
    neg, _ = s.factorize(K, v)
    st = s.stats()
    assert st["lu_fallbacks"] == 0 and neg > 0 and st["probe_backward_error_digits"] >= 13
