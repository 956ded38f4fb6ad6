"""Pin the CPU oracle (oracle/nr_oracle.c) before trusting it.

Each check is chosen so that a plausible mistake in the restatement fails it:
  * known answers: the reference's 4-bus smoke grid and a 30-bus grid, against an
    independent dense scipy Newton (tests/golden/newton_small.npz);
  * closed forms: the oracle's Jacobian equals the textbook dS/dV Jacobian and a finite
    difference of the mismatch (a dropped term, sign or index fails);
  * invariants of the reference's own tests: thread invariance, batch column == fresh
    single solve with a diverging column in the middle, NaN start voltage raises, NaN
    admittance converges nowhere, pinned buses stay frozen;
  * the pivoting LU standing in for KLU equals a dense solve, including matrices that
    need row pivoting.
"""

import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle
from oracle.oracle import OracleSolver
from p3smi import grids

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "newton_small.npz"))


def csr(Y):
    M = sp.csr_matrix(Y)
    M.sort_indices()
    return M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.complex128)


def mismatch(Y, V, S, pv, pq):
    mis = V * np.conj(Y @ V) - S
    return float(np.max(np.abs(np.r_[mis[np.r_[pv, pq]].real, mis[pq].imag])))


def textbook_jacobian(Y, V, pv, pq):
    """dS/dV polar Jacobian (p3s/PQPVPowerflow.py:181-189 form), dense."""
    Y = np.asarray(Y)
    pvpq = np.r_[pv, pq]
    Ibus = Y @ V
    dV = np.diag(V)
    dVn = np.diag(V / np.abs(V))
    dS_dVm = dV @ np.conj(Y @ dVn) + np.conj(np.diag(Ibus)) @ dVn
    dS_dVa = 1j * dV @ np.conj(np.diag(Ibus) - Y @ dV)
    return np.block([[dS_dVa[np.ix_(pvpq, pvpq)].real, dS_dVm[np.ix_(pvpq, pq)].real],
                     [dS_dVa[np.ix_(pq, pvpq)].imag, dS_dVm[np.ix_(pq, pq)].imag]])


def test_four_bus_known_answer():
    Y, pv, pq, S, V0 = (GOLD[k] for k in ("four_Y", "four_pv", "four_pq", "four_S", "four_V0"))
    Yp, Yj, Yx = csr(Y)
    r = oracle.solve_single(Yp, Yj, Yx, S, V0, pv, pq, max_iter=30, tol=1e-10)
    assert r["converged"]
    assert mismatch(Y, r["V"], S, pv, pq) < 1e-9
    np.testing.assert_allclose(r["V"], GOLD["four_V"], atol=1e-11)
    assert r["iterations"] == int(GOLD["four_it"])
    s = OracleSolver(Yp, Yj, Yx, pv, pq)
    rs = s.solve(S, V0, max_iter=30, tol=1e-10)
    np.testing.assert_allclose(rs["V"], r["V"], atol=1e-10)
    rb1 = s.solve_batch(GOLD["four_Sb"], V0, max_iter=30, tol=1e-10, n_threads=1)
    rb4 = s.solve_batch(GOLD["four_Sb"], V0, max_iter=30, tol=1e-10, n_threads=4)
    assert rb1["converged"].all() and rb4["converged"].all()
    np.testing.assert_array_equal(rb1["V"], rb4["V"])
    np.testing.assert_allclose(rb1["V"], GOLD["four_Vb"], atol=1e-11)
    np.testing.assert_array_equal(rb1["iterations"], GOLD["four_itb"])
    worst = max(mismatch(Y, rb4["V"][:, k], GOLD["four_Sb"][:, k], pv, pq) for k in range(16))
    assert worst < 1e-9


def test_geo30_matches_dense_newton():
    Y, pv, pq, S, V0 = (GOLD[k] for k in ("geo30_Y", "geo30_pv", "geo30_pq", "geo30_S", "geo30_V0"))
    Yp, Yj, Yx = csr(Y)
    r = OracleSolver(Yp, Yj, Yx, pv, pq).solve_batch(S, V0, max_iter=20, tol=1e-8)
    assert r["converged"].all()
    np.testing.assert_array_equal(r["iterations"], GOLD["geo30_it"])
    np.testing.assert_allclose(r["V"], GOLD["geo30_V"], atol=1e-10)


@pytest.mark.parametrize("name", ["pf200"])
def test_jacobian_closed_form_and_finite_difference(name):
    g = grids.config_grid(name, 0)
    Yp, Yj, Yx = g.csr()
    rng = np.random.default_rng(3)
    V = g.flat_start() * (1 + 0.05 * rng.standard_normal(g.n)) * np.exp(1j * 0.2 * rng.standard_normal(g.n))
    s = OracleSolver(Yp, Yj, Yx, g.pv, g.pq)
    d = s.debug_J_at(g.sbus(), V)
    J = sp.csc_matrix((d["Jx"], d["Ji"], d["Jp"]), shape=(d["m"], d["m"])).toarray()
    Jt = textbook_jacobian(g.ybus().toarray(), V, g.pv, g.pq)
    assert np.abs(J - Jt).max() <= 1e-12 * np.abs(Jt).max()
    # F = Scalc - Sspec, finite difference of F along the Newton variables
    Y = g.ybus().toarray()
    pvpq = np.r_[g.pv, g.pq]

    def F(Va, Vm):
        Vv = Vm * np.exp(1j * Va)
        mis = Vv * np.conj(Y @ Vv) - g.sbus()
        return np.r_[mis[pvpq].real, mis[g.pq].imag]

    np.testing.assert_allclose(d["F"], F(np.angle(V), np.abs(V)), atol=1e-12)
    h = 1e-7
    for col in rng.choice(d["m"], 12, replace=False):
        Va, Vm = np.angle(V).copy(), np.abs(V).copy()
        if col < len(pvpq):
            Va[pvpq[col]] += h
        else:
            Vm[g.pq[col - len(pvpq)]] += h
        fd = (F(Va, Vm) - d["F"]) / h
        assert np.abs(fd - J[:, col]).max() <= 1e-5 * max(1.0, np.abs(J[:, col]).max())


def test_batch_column_equals_fresh_single_with_diverging_column():
    """tests/cpp/test_cpp_timeseries.py:108-144: every batch column equals a fresh single
    solve (same flag, same iterations, |dV| < 1e-12), a x5-load column in the middle."""
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    S = grids.timeseries_scaled(g, 9, seed=2)
    S[:, 4] *= 5.0
    s = OracleSolver(Yp, Yj, Yx, g.pv, g.pq)
    rb = s.solve_batch(S, g.flat_start(), max_iter=20, tol=1e-8, n_threads=3)
    assert not rb["converged"][4]
    for k in range(9):
        r1 = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve(S[:, k], g.flat_start(), max_iter=20, tol=1e-8)
        assert r1["converged"] == rb["converged"][k]
        assert r1["iterations"] == rb["iterations"][k]
        if r1["converged"]:
            assert np.abs(r1["V"] - rb["V"][:, k]).max() < 1e-12


def test_thread_invariance_bit_identical():
    g = grids.config_grid("ts2000", 0)
    Yp, Yj, Yx = g.csr()
    S = grids.timeseries_daily(g, 48, seed=1)
    s = OracleSolver(Yp, Yj, Yx, g.pv, g.pq)
    r1 = s.solve_batch(S, g.flat_start(), max_iter=20, tol=1e-8, n_threads=1)
    r8 = s.solve_batch(S, g.flat_start(), max_iter=20, tol=1e-8, n_threads=8)
    np.testing.assert_array_equal(r1["V"], r8["V"])
    np.testing.assert_array_equal(r1["iterations"], r8["iterations"])
    assert s.lean_active


def test_nan_start_voltage_raises():
    """tests/cpp/test_cpp_threading_guards.py:62-69."""
    Y, pv, pq, S, V0 = (GOLD[k] for k in ("four_Y", "four_pv", "four_pq", "four_S", "four_V0"))
    s = OracleSolver(*csr(Y), pv, pq)
    V0 = V0.copy()
    V0[2] = np.nan
    with pytest.raises(RuntimeError, match="solve_batch: invalid initial voltage at bus 2"):
        s.solve_batch(S.reshape(-1, 1), V0)


def test_nan_admittance_converges_nowhere():
    """tests/cpp/test_cpp_threading_guards.py:72-85: NaN on a PQ bus's Ybus diagonal
    (set through update_Y) never converges; a NaN in the slack row is harmless."""
    Y, pv, pq, S, V0 = (GOLD[k] for k in ("four_Y", "four_pv", "four_pq", "four_S", "four_V0"))
    Yp, Yj, Yx = csr(Y)
    s = OracleSolver(Yp, Yj, Yx, pv, pq)
    row = int(pq[0])
    k = Yp[row] + int(np.flatnonzero(Yj[Yp[row]:Yp[row + 1]] == row)[0])
    bad = Yx.copy()
    bad[k] = complex(np.nan, np.nan)
    s.update_Y(bad)
    r = s.solve_batch(GOLD["four_Sb"], V0, max_iter=10, tol=1e-8)
    assert not r["converged"].any()
    slack_row = Yx.copy()
    slack_row[Yp[0] + 1] = np.nan
    s.update_Y(slack_row)
    assert s.solve_batch(GOLD["four_Sb"], V0, max_iter=10, tol=1e-8)["converged"].all()


def test_pinned_buses_stay_frozen():
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    L = 3
    Ym = np.repeat(np.abs(Yx)[:, None], L, axis=1)
    Ya = np.repeat(np.angle(Yx)[:, None], L, axis=1)
    V0 = np.repeat(g.flat_start()[:, None], L, axis=1)
    pin = np.zeros((g.n, L), np.uint8)
    pq0 = g.pq[:5]
    pin[pq0, 1] = 1
    V0[pq0, 1] = 0.97 * np.exp(0.05j)
    r = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch_contingency(Ym, Ya, g.sbus(), V0, pin, 20, 1e-8)
    assert r["converged"].all()
    np.testing.assert_allclose(r["V"][pq0, 1], V0[pq0, 1], rtol=0, atol=1e-15)
    # an unpinned column equals the time-series solve (|Y|, angle(Y) rounded by numpy here)
    rt = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve(g.sbus(), g.flat_start(), 20, 1e-8)
    np.testing.assert_allclose(r["V"][:, 0], rt["V"], rtol=0, atol=1e-12)
    assert r["iterations"][0] == rt["iterations"]


def test_pivoting_lu_matches_dense_solve():
    rng = np.random.default_rng(0)
    for m in (5, 40, 120):
        A = sp.random(m, m, density=0.08, random_state=int(rng.integers(1 << 30))).toarray()
        A += np.diag(rng.uniform(0.5, 2.0, m))
        # force row pivoting: zero a few diagonal entries, add a matching off-diagonal
        for i in rng.choice(m, max(1, m // 5), replace=False):
            A[i, i] = 0.0
            A[(i + 1) % m, i] = 3.0
        if abs(np.linalg.det(A)) < 1e-8:
            continue
        b = rng.standard_normal(m)
        for ordering in (0, 2):
            x = oracle.gp_solve(sp.csc_matrix(A), b, ordering=ordering)
            np.testing.assert_allclose(A @ x, b, atol=1e-9 * max(1, np.abs(b).max()))


def test_minimum_degree_is_a_permutation_and_reduces_fill():
    n = 20
    T = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(n, n))
    A = (sp.kron(sp.identity(n), T) + sp.kron(T, sp.identity(n))).tocsc()
    p = oracle.min_degree(A)
    assert sorted(p.tolist()) == list(range(n * n))

    def fill(perm):
        from scipy.sparse.linalg import splu

        M = sp.csc_matrix(A.toarray()[np.ix_(perm, perm)])
        lu = splu(M, permc_spec="NATURAL", diag_pivot_thresh=0.0, options={"SymmetricMode": True})
        return int(lu.L.nnz + lu.U.nnz)

    assert fill(p) < fill(np.arange(n * n))


@pytest.mark.parametrize("tag,max_iter,ls", [("ls", 20, True), ("ls_lim", 5, True), ("nols", 20, False),
                                              ("nols_lim", 5, False)])
def test_damped_newton_matches_independent_dense_restatement(tag, max_iter, ls):
    """Golden vectors from a dense numpy Newton that restates the reference's damping rules
    (tests/golden/make_golden.py:newton_armijo_dense, p3s/cpp/nr_klu.cpp:588-708) on 48
    perturbed starts: Armijo backtracking (12 cases), the iteration limit (max_iter 5) and,
    without line search, divergence. Flags and iteration counts are exact; voltages of the
    converged cases and of the backtracking cases stopped at the limit agree to roundoff
    (diverging full-step trajectories amplify roundoff and are not compared)."""
    g = GOLD
    Y = sp.csr_matrix(g["geo30_Y"])
    Y.sort_indices()
    o = OracleSolver(Y.indptr, Y.indices, Y.data, g["geo30_pv"], g["geo30_pq"])
    r = o.solve_batch(g["damped_S"], g["damped_V0"], max_iter=max_iter, tol=1e-8, line_search=ls)
    cv, it, V = g[f"damped_{tag}_cv"], g[f"damped_{tag}_it"], g[f"damped_{tag}_V"]
    np.testing.assert_array_equal(r["converged"], cv)
    np.testing.assert_array_equal(r["iterations"], it)
    keep = cv if not ls else np.ones_like(cv)
    assert np.abs(r["V"][:, keep] - V[:, keep]).max() < 1e-12


def test_best_trial_commit_matches_independent_dense_restatement():
    """Loads past the transfer limit: the line search finds no sufficient decrease and commits
    its least-bad trial (p3s/cpp/nr_klu.cpp:688-697). After 6 iterations (best-trial commits
    in most cases) the state equals the dense restatement's to roundoff; after 20, flags and
    iteration counts (the limit) are exact."""
    g = GOLD
    Y = sp.csr_matrix(g["geo30_Y"])
    Y.sort_indices()
    o = OracleSolver(Y.indptr, Y.indices, Y.data, g["geo30_pv"], g["geo30_pq"])
    r6 = o.solve_batch(g["overload_S"], g["geo30_V0"], max_iter=6, tol=1e-8, line_search=True)
    np.testing.assert_array_equal(r6["converged"], g["overload6_cv"])
    np.testing.assert_array_equal(r6["iterations"], g["overload6_it"])
    assert np.abs(r6["V"] - g["overload6_V"]).max() < 1e-10
    r20 = o.solve_batch(g["overload_S"], g["geo30_V0"], max_iter=20, tol=1e-8, line_search=True)
    np.testing.assert_array_equal(r20["converged"], g["overload20_cv"])
    np.testing.assert_array_equal(r20["iterations"], g["overload20_it"])


def test_rerun_on_pivoting_factorization_equals_pivoting_solve():
    """run_column (p3s/cpp/nr_klu.cpp:844-857): a column that does not converge on the frozen
    pivot order is solved again from scratch on a pivoting factorization. So every
    non-converged column must equal, bit for bit, the solve in which the frozen plan is
    always rejected (every factorization pivoting from the first iteration)."""
    om = oracle
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    rng = np.random.default_rng(7)
    T = 128
    k = rng.uniform(1.0, 2.5, T)
    S = g.p_gen[:, None] * k[None, :] - (g.p_load + 1j * g.q_load)[:, None] * k[None, :]
    V0 = np.repeat(g.flat_start()[:, None], T, axis=1)
    V0[g.pq, :] *= rng.uniform(0.6, 1.4, T)[None, :]
    o = OracleSolver(Yp, Yj, Yx, g.pv, g.pq)
    try:
        om.reset_counters()
        r0 = o.solve_batch(S, V0, max_iter=20, tol=1e-8)
        reruns = om.counters()["column_reruns"]
        om.set_lean_guard(1e300)  # the frozen plan never passes: pivoting factorizations only
        r1 = o.solve_batch(S, V0, max_iter=20, tol=1e-8)
    finally:
        om.set_lean_guard(1e-3)
    bad = ~r0["converged"]
    assert reruns == int(bad.sum()) > 0
    np.testing.assert_array_equal(r1["converged"][bad], r0["converged"][bad])
    np.testing.assert_array_equal(r1["iterations"][bad], r0["iterations"][bad])
    np.testing.assert_array_equal(np.ascontiguousarray(r1["V"][:, bad]).view(np.float64),
                                  np.ascontiguousarray(r0["V"][:, bad]).view(np.float64))
