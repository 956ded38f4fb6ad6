"""Uneven batching: batches of several topologies solved together (p3smi_group_*).

Grids of different sizes (and LU launch shapes) with different case counts run every Newton
pass as one launch per kernel over all of them, the passes decided on the device and captured
several to a graph launch. Every job's result must be bit-identical to solving that job alone
(flags, iteration counts and voltage bits): every LU launch shape sums in one order, and no
arithmetic mixes batches. Replays of the captured graph are checked against eager solves run
in between. A sample per size bucket meets the oracle bar (flags and iteration counts exact,
|dV| <= 1e-9). A member whose cases need pivoting factorizations (guard forced, as the
reference's tests/cuda/test_lean_lu.py:224-244 does) and a member with line searches and
reruns (heavy loads from perturbed starts) rejoin the group after their host-side passes."""

import numpy as np
import pytest

from oracle import oracle as ora
from oracle.oracle import OracleSolver
from p3smi import autotune, grids, nr_klu
from p3smi.group import Group, solve_group
from p3smi.timeseries import dc_initial_voltage

pytestmark = pytest.mark.gpu

TOL_V = 1e-9


def _same(res, alone, tag):
    np.testing.assert_array_equal(res["converged"], alone["converged"], err_msg=tag)
    np.testing.assert_array_equal(res["iterations"], alone["iterations"], err_msg=tag)
    np.testing.assert_array_equal(res["V"].view(np.float64), alone["V"].view(np.float64), err_msg=tag)


def _counts(rng, k, lo, hi):
    return np.clip(np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), k))), lo, hi).astype(int)


def _family_jobs(count, lo, hi, seed, cmax=256):
    rng = np.random.default_rng(seed)
    fam = grids.grid_family(count, lo, hi, seed=seed)
    counts = _counts(rng, count, 1, cmax)
    jobs = []
    for k, (g, b) in enumerate(zip(fam, counts)):
        Yp, Yj, Yx = g.csr()
        s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
        autotune.apply_cached(s, int(b))  # each member in its own tuned LU launch shape
        jobs.append((g, s, np.ascontiguousarray(grids.timeseries_scaled(g, int(b), seed=200 + k, sigma=0.1)),
                     np.ascontiguousarray(dc_initial_voltage(g))))
    return jobs


def test_group_of_uneven_grids_equals_alone_and_replays(gpu):
    jobs = _family_jobs(16, 100, 8000, seed=9)
    shapes = {tuple(s.info[k] for k in ("lu_g", "lu_ns", "lu_few_g", "lu_few_ns")) for _, s, _, _ in jobs}
    print(f"group: {len(jobs)} grids ({min(g.n for g, *_ in jobs)}..{max(g.n for g, *_ in jobs)} buses), "
          f"{sum(S.shape[1] for _, _, S, _ in jobs)} cases, LU shapes {sorted(shapes)}")
    grp = Group([(s, S, V0) for _, s, S, V0 in jobs])
    try:
        for rnd in range(3):
            res = grp.solve(max_iter=20, tol=1e-8, passes_per_graph=8)
            print(f"round {rnd}: passes {grp.passes()}")
            for k, ((g, s, S, V0), r) in enumerate(zip(jobs, res)):
                alone = s.solve_batch(S, V0, max_iter=20, tol=1e-8)  # eager work between the replays
                _same(r, alone, f"round {rnd} grid {k} ({g.n} buses, {S.shape[1]} cases)")
    finally:
        grp.close()
    checked = set()
    for k, ((g, s, S, V0), r) in enumerate(zip(jobs, res)):
        bk = 0 if g.n <= 500 else (1 if g.n <= 5000 else 2)
        if bk in checked:
            continue
        checked.add(bk)
        cols = np.unique(np.linspace(0, S.shape[1] - 1, min(8, S.shape[1])).round().astype(int))
        Yp, Yj, Yx = g.csr()
        ref = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(np.ascontiguousarray(S[:, cols]), V0, 20, 1e-8)
        np.testing.assert_array_equal(r["converged"][cols], ref["converged"])
        np.testing.assert_array_equal(r["iterations"][cols], ref["iterations"])
        ok = ref["converged"]
        Vg = r["V"][:, cols][:, ok]
        assert np.abs(np.abs(Vg) - np.abs(ref["V"][:, ok])).max(initial=0.0) <= TOL_V
        assert np.abs(np.angle(Vg) - np.angle(ref["V"][:, ok])).max(initial=0.0) <= TOL_V
    assert checked == {0, 1, 2}


def test_passes_per_graph_do_not_change_the_bits(gpu):
    jobs = _family_jobs(6, 150, 3000, seed=11, cmax=96)
    first = None
    for K in (1, 3, 8):
        res = solve_group([(s, S, V0) for _, s, S, V0 in jobs], max_iter=20, tol=1e-8, passes_per_graph=K)
        if first is None:
            first = res
        else:
            for k, (a, b) in enumerate(zip(res, first)):
                _same(a, b, f"K={K} grid {k}")


def test_group_member_with_pivoting_and_reruns(gpu):
    """One member has every frozen factorization rejected (guard 1e30), one has heavy loads
    from perturbed starts (line searches, non-converging cases rerun with pivoting); both take
    their pivoting passes on the host side and rejoin. Results equal each solved alone, and
    the heavy member's flags and counts equal the oracle's."""
    g200 = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g200.csr()
    rng = np.random.default_rng(7)
    T = 128
    k = rng.uniform(1.0, 2.5, T)
    S_heavy = g200.p_gen[:, None] * k[None, :] - (g200.p_load + 1j * g200.q_load)[:, None] * k[None, :]
    V_heavy = np.repeat(g200.flat_start()[:, None], T, axis=1)
    V_heavy[g200.pq, :] *= rng.uniform(0.6, 1.4, T)[None, :]
    heavy = nr_klu.Solver(Yp, Yj, Yx, g200.pv, g200.pq)
    g2k = grids.config_grid("ts2000", 0)
    forced = nr_klu.Solver(*g2k.csr(), g2k.pv, g2k.pq)
    forced.set_guard(1e30)
    S2k = np.ascontiguousarray(grids.timeseries_daily(g2k, 8760, seed=1)[:, ::97])
    V2k = np.ascontiguousarray(dc_initial_voltage(g2k))
    rest = _family_jobs(4, 300, 6000, seed=13, cmax=64)
    jobs = [(heavy, np.ascontiguousarray(S_heavy), np.ascontiguousarray(V_heavy)), (forced, S2k, V2k)]
    jobs += [(s, S, V0) for _, s, S, V0 in rest]
    grp = Group(jobs)
    try:
        res = grp.solve(max_iter=20, tol=1e-8)
        passes = grp.passes()
        stats = grp.stats()
    finally:
        grp.close()
    print(f"passes {passes}; heavy: reruns {stats[0]['reruns']}, pivoting {stats[0]['pivot_factorizations']}; "
          f"forced: guard trips {stats[1]['guard_trips']}, pivoting {stats[1]['pivot_factorizations']}")
    assert passes["host"] > 0
    assert stats[0]["reruns"] > 0 and stats[1]["guard_trips"] == S2k.shape[1]
    for i, ((s, S, V0), r) in enumerate(zip(jobs, res)):
        _same(r, s.solve_batch(S, V0, max_iter=20, tol=1e-8), f"member {i}")
    ora.reset_counters()
    ref = OracleSolver(Yp, Yj, Yx, g200.pv, g200.pq).solve_batch(S_heavy, V_heavy, max_iter=20, tol=1e-8)
    np.testing.assert_array_equal(res[0]["converged"], ref["converged"])
    np.testing.assert_array_equal(res[0]["iterations"], ref["iterations"])
    assert stats[0]["reruns"] == ora.counters()["column_reruns"]
