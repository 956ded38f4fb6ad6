"""GPU == CPU oracle on the same inputs: identical convergence flags and iteration counts,
max |dVm| <= 1e-9 p.u. and max |dVa| <= 1e-9 rad (the reference's own GPU test demands
the same of its GPU path against nr_klu, tests/cuda/test_lean_lu.py:126-147)."""

import numpy as np
import pytest

from oracle.oracle import OracleSolver
from p3smi import grids, nr_klu
from p3smi.timeseries import dc_initial_voltage

pytestmark = pytest.mark.gpu

TOL_V = 1e-9


def _compare(gpu, ref, tag=""):
    assert np.array_equal(gpu["converged"], ref["converged"]), f"{tag}: converged flags differ"
    assert np.array_equal(gpu["iterations"], ref["iterations"]), (
        f"{tag}: iterations differ at {np.nonzero(gpu['iterations'] != ref['iterations'])[0][:10]}")
    ok = ref["converged"]
    dvm = np.abs(np.abs(gpu["V"][:, ok]) - np.abs(ref["V"][:, ok])).max(initial=0.0)
    dva = np.abs(np.angle(gpu["V"][:, ok]) - np.angle(ref["V"][:, ok])).max(initial=0.0)
    assert dvm <= TOL_V, f"{tag}: max|dVm| = {dvm:.3e}"
    assert dva <= TOL_V, f"{tag}: max|dVa| = {dva:.3e}"
    return dvm, dva


def test_pf200_timeseries(gpu):
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    S = grids.timeseries_scaled(g, 96, seed=1)
    V0 = g.flat_start()
    r = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8)
    ref = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8)
    _compare(r, ref, "pf200")
    assert r["converged"].all()
    assert (r["mismatch"] < 1e-8).all()


def test_ts2000_subset_dc_start(gpu):
    g = grids.config_grid("ts2000", 0)
    Yp, Yj, Yx = g.csr()
    S = grids.timeseries_daily(g, 8760, seed=1)[:, ::17].copy()  # 516 hours spread over the year
    V0 = dc_initial_voltage(g)
    r = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8)
    ref = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8)
    _compare(r, ref, "ts2000")


def test_line_search_from_perturbed_starts(gpu):
    """Heavier loads (x1..2.5) and PQ magnitudes started at x0.6..1.4 make full steps
    overshoot: the backtracking decisions, the pivot-ratio fallback and the iteration
    limit must all land as in the oracle (tests/cuda/test_lean_lu.py:150-174)."""
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    rng = np.random.default_rng(7)
    T = 256
    k = rng.uniform(1.0, 2.5, T)
    S = g.p_gen[:, None] * k[None, :] - (g.p_load + 1j * g.q_load)[:, None] * k[None, :]
    V0 = np.repeat(g.flat_start()[:, None], T, axis=1)
    V0[g.pq, :] *= rng.uniform(0.6, 1.4, T)[None, :]
    for ls in (True, False):
        r = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8, line_search=ls)
        ref = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8, line_search=ls)
        _compare(r, ref, f"perturbed ls={ls}")
        assert 0 < r["converged"].sum() < T


def test_nan_column_fails_alone(gpu):
    """A NaN injection in one column fails that column only (tests/cuda/test_lean_lu.py:212-221)."""
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    S = grids.timeseries_scaled(g, 70, seed=3)
    S[5, 13] = np.nan
    V0 = g.flat_start()
    r = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8)
    ref = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8)
    assert not r["converged"][13]
    assert r["converged"][np.arange(70) != 13].all()
    assert r["iterations"][13] == ref["iterations"][13] == 0
    keep = np.arange(70) != 13
    _compare({k: (v[:, keep] if k == "V" else v[keep]) for k, v in r.items()},
             {k: (v[:, keep] if k == "V" else v[keep]) for k, v in ref.items()}, "nan column")


@pytest.mark.parametrize("name", ["pf200", "ts2000"])
def test_mixed_precision_matches_fp64_oracle(gpu, name):
    """FP32 factors + FP64 refinement: same flags and iterations as the FP64 oracle and
    |dV| <= 1e-9 (the refined step is FP64-accurate)."""
    g = grids.config_grid(name, 0)
    Yp, Yj, Yx = g.csr()
    if name == "pf200":
        S = grids.timeseries_scaled(g, 96, seed=1)
        V0 = g.flat_start()
    else:
        S = grids.timeseries_daily(g, 8760, seed=1)[:, ::23].copy()
        V0 = dc_initial_voltage(g)
    r = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8, precision="mixed")
    ref = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8)
    _compare(r, ref, f"{name} mixed")


def test_mixed_precision_perturbed_starts(gpu):
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    rng = np.random.default_rng(7)
    T = 256
    k = rng.uniform(1.0, 2.5, T)
    S = g.p_gen[:, None] * k[None, :] - (g.p_load + 1j * g.q_load)[:, None] * k[None, :]
    V0 = np.repeat(g.flat_start()[:, None], T, axis=1)
    V0[g.pq, :] *= rng.uniform(0.6, 1.4, T)[None, :]
    r = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8, precision="mixed")
    ref = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8)
    _compare(r, ref, "perturbed mixed")


# (cases per workgroup, wavefronts, group): group 0 at 64 cases is the 2x2-block schedule
# (factorization and substitutions), group 1 the one-entry-per-task schedule of the same shape
SHAPES = [(64, 16, 0), (64, 16, 1), (64, 8, 0), (64, 8, 1), (32, 16, 0), (32, 8, 0), (16, 16, 0), (16, 8, 0),
          (1, 16, 0), (1, 8, 0)]


@pytest.mark.parametrize("precision", ["double", "mixed", "mixed_switch"])
@pytest.mark.parametrize("name", ["pf200", "ts2000"])
def test_every_lu_shape_gives_the_same_bits(gpu, precision, name):
    """Every launch shape sums each factor entry in one order, so a case's result does not
    depend on the shape that solved it (serving relies on this); the results also meet the
    oracle bar."""
    g = grids.config_grid(name, 0)
    Yp, Yj, Yx = g.csr()
    if name == "pf200":
        S, V0 = grids.timeseries_scaled(g, 192, seed=3), g.flat_start()
    else:
        S, V0 = grids.timeseries_daily(g, 128, seed=3), dc_initial_voltage(g)
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    first = None
    for gg, ns, grp in SHAPES:
        s.set_lu_shape(gg, ns, 1, 16, 0, grp)  # few_max 0: every launch uses the batch shape
        r = s.solve_batch(S, V0, max_iter=20, tol=1e-8, precision=precision)
        if first is None:
            first = r
            if precision == "double":
                ref = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8)
                _compare(r, ref, f"shape {gg}x{ns}")
            continue
        assert np.array_equal(r["iterations"], first["iterations"]), (gg, ns, grp)
        assert np.array_equal(r["V"].view(np.float64), first["V"].view(np.float64)), (gg, ns, grp)
        assert np.array_equal(r["mismatch"], first["mismatch"]), (gg, ns, grp)


def test_block_lu_group_width_gives_the_same_bits(gpu):
    """The FP64 block LU runs 48 cases per workgroup while a batch's groups of 48 fit one per
    compute unit (a 512-case batch on 256 compute units), 64 otherwise (12,800 cases). The
    same cases give the same bits either way, wherever they sit in the batch."""
    g = grids.config_grid("ts2000", 0)
    Yp, Yj, Yx = g.csr()
    S = grids.timeseries_daily(g, 512, seed=5)
    V0 = dc_initial_voltage(g)
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    s.set_lu_shape(64, 16, 1, 16, 0, 0)  # few_max 0: every launch uses the batch shape
    small = s.solve_batch(S, V0, max_iter=20, tol=1e-8)
    big = s.solve_batch(np.tile(S, (1, 25)), V0, max_iter=20, tol=1e-8)
    for k in (0, 7, 24):
        part = slice(512 * k, 512 * (k + 1))
        assert np.array_equal(big["iterations"][part], small["iterations"]), k
        assert np.array_equal(big["V"][:, part].view(np.float64), small["V"].view(np.float64)), k
        assert np.array_equal(big["mismatch"][part], small["mismatch"]), k


def test_switching_mixed_mode(gpu):
    """The reference GPU solver's mixed rule (p3s/cuda/nr_polar_kernels.cu:209-288): FP32 steps
    while the mismatch is above 1e-3 p.u., FP64 after. The converged voltages meet the 1e-9 bar
    against the FP64 oracle, and flags and iteration counts are exact: far from the solution
    an FP32 step's relative error (~cond(J) x 6e-8) is far below the Newton contraction, and
    the last steps, below 1e-3, are FP64 (the reference only guarantees +1 iteration,
    tests/cuda/test_lean_lu.py:201; here equality holds on these sets and is required)."""
    for name, T in (("pf200", 96), ("ts2000", 512)):
        g = grids.config_grid(name, 0)
        Yp, Yj, Yx = g.csr()
        if name == "pf200":
            S, V0 = grids.timeseries_scaled(g, T, seed=1), g.flat_start()
        else:
            S, V0 = grids.timeseries_daily(g, T, seed=1), dc_initial_voltage(g)
        solver = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
        r = solver.solve_batch(S, V0, max_iter=20, tol=1e-8, precision="mixed_switch")
        ref = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8)
        assert np.array_equal(r["converged"], ref["converged"]), name
        assert np.array_equal(r["iterations"], ref["iterations"]), (
            name, np.nonzero(r["iterations"] != ref["iterations"])[0][:10])
        ok = ref["converged"]
        dvm = np.abs(np.abs(r["V"][:, ok]) - np.abs(ref["V"][:, ok])).max()
        dva = np.abs(np.angle(r["V"][:, ok]) - np.angle(ref["V"][:, ok])).max()
        assert dvm <= TOL_V and dva <= TOL_V, (name, dvm, dva)
        assert (r["mismatch"][ok] < 1e-8).all()
        # one factorization per case and iteration; FP32 pivot failures add an FP64 one
        st = solver.last_stats
        assert st["single_factorizations"] > 0, st
        assert st["single_factorizations"] + st["double_factorizations"] >= r["iterations"].sum(), st
        print(f"{name}: {st['single_factorizations']} FP32 / {st['double_factorizations']} FP64 factorizations")


def test_last_stats_count_one_factorization_per_iteration(gpu):
    """FP64: every active case is factored once per Newton iteration, so the solve's FP64
    factorization count equals the sum of the iteration counts (also over a call that the
    workspace limit splits: the counters of every chunk are summed)."""
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    S = grids.timeseries_scaled(g, 300, seed=5)
    solver = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    r = solver.solve_batch(S, g.flat_start(), max_iter=20, tol=1e-8)
    st = solver.last_stats
    assert st["cases"] == 300
    assert st["double_factorizations"] == r["iterations"].sum(), st
    assert st["single_factorizations"] == 0
    # the same call in chunks of at most 128 cases (128, 128, 44): identical bits, summed counters
    solver.set_max_chunk(128)
    rc = solver.solve_batch(S, g.flat_start(), max_iter=20, tol=1e-8)
    solver.set_max_chunk(0)
    assert np.array_equal(rc["V"].view(np.float64), r["V"].view(np.float64))
    assert np.array_equal(rc["iterations"], r["iterations"])
    assert np.array_equal(rc["mismatch"], r["mismatch"])
    assert solver.last_stats["cases"] == 300
    assert solver.last_stats["double_factorizations"] == st["double_factorizations"]


def test_update_Y_equals_a_rebuilt_solver(gpu):
    """Solver.update_Y (p3s/cpp/nr_klu.cpp:999-1011) refreshes the admittances on the device
    without a new symbolic set-up: solving afterwards gives the bits a solver built from the
    new values gives (tests/cuda/test_lean_lu.py:270-284), and a wrong nnz is refused."""
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    S = grids.timeseries_scaled(g, 64, seed=2)
    V0 = g.flat_start()
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    first = s.solve_batch(S, V0, max_iter=20, tol=1e-8)  # device buffers resident before the update
    rng = np.random.default_rng(3)
    Yx2 = Yx * (1.0 + 0.03 * rng.uniform(-1.0, 1.0, Yx.shape))  # same pattern, new values
    s.update_Y(Yx2)
    upd = s.solve_batch(S, V0, max_iter=20, tol=1e-8)
    fresh = nr_klu.Solver(Yp, Yj, Yx2, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8)
    np.testing.assert_array_equal(upd["iterations"], fresh["iterations"])
    np.testing.assert_array_equal(upd["converged"], fresh["converged"])
    np.testing.assert_array_equal(upd["V"].view(np.float64), fresh["V"].view(np.float64))
    assert not np.array_equal(upd["V"], first["V"])
    ref = OracleSolver(Yp, Yj, Yx2, g.pv, g.pq).solve_batch(S, V0, max_iter=20, tol=1e-8)
    _compare(upd, ref, "update_Y")
    with pytest.raises(Exception, match="nnz"):
        s.update_Y(Yx2[:-1])


def test_evaluation_tiles_do_not_change_the_bits(gpu):
    """The mismatch/Jacobian evaluation's row tiles (rows per workgroup, staged in LDS up to 768
    nonzeros) only change which workgroup evaluates a row: every shape gives the same bits."""
    g = grids.config_grid("ts2000", 0)
    Yp, Yj, Yx = g.csr()
    S, V0 = grids.timeseries_daily(g, 128, seed=5), dc_initial_voltage(g)
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    first = None
    for rows in (32, 1, 8, 64, 1024):
        s.set_eval_tile(rows)
        assert s.info["eval_rows"] == rows
        r = s.solve_batch(S, V0, max_iter=20, tol=1e-8)
        if first is None:
            first = r
            continue
        np.testing.assert_array_equal(r["iterations"], first["iterations"])
        np.testing.assert_array_equal(r["V"].view(np.float64), first["V"].view(np.float64))
        np.testing.assert_array_equal(r["mismatch"], first["mismatch"])
    with pytest.raises(Exception, match="rows"):
        s.set_eval_tile(0)
