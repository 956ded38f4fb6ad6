"""Continuous batching at the scale of the multi-grid and serving workloads.

multi48: 48 synthetic grids (100..20,000 buses, log-uniform), 4,096 load-scaled cases in
batches of 1..256 per grid, all submitted at once to the server. Poisson mix: requests arriving
at 200/s over 8 topologies up to 25,000 buses, 70 % time series and 30 % N-1 subsets. Every
request's result must be bit-identical to solving it alone (slots never mix arithmetic and
every LU launch shape sums in one order), and a sample of cases per size bucket (<= 500,
<= 5k, <= 20k buses, and the 25k grid) meets the oracle bar: flags and iteration counts exact,
|dV| <= 1e-9."""

import time

import numpy as np
import pytest

from oracle.oracle import OracleSolver
from p3smi import contingency, grids, nr_klu
from p3smi.serving import Server, topology_key
from p3smi.timeseries import dc_initial_voltage

pytestmark = pytest.mark.gpu

TOL_V = 1e-9


def _counts(rng, k, total, lo, hi):
    """k batch sizes, log-uniform in [lo, hi], rescaled to sum to `total`."""
    c = np.exp(rng.uniform(np.log(lo), np.log(hi), k))
    c = np.clip(np.round(c * total / c.sum()), lo, hi).astype(int)
    while c.sum() != total:
        d = total - c.sum()
        i = int(np.argmax(c)) if d < 0 else int(np.argmin(c))
        c[i] = int(np.clip(c[i] + np.sign(d) * min(abs(d), hi), lo, hi))
    return c


def _same(res, alone, tag):
    np.testing.assert_array_equal(res["converged"], alone["converged"], err_msg=tag)
    np.testing.assert_array_equal(res["iterations"], alone["iterations"], err_msg=tag)
    np.testing.assert_array_equal(res["V"].view(np.float64), alone["V"].view(np.float64), err_msg=tag)


def _oracle_check(g, S, V0, res, cols, tag):
    Yp, Yj, Yx = g.csr()
    ref = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(np.ascontiguousarray(S[:, cols]), V0, 20, 1e-8)
    np.testing.assert_array_equal(res["converged"][cols], ref["converged"], err_msg=tag)
    np.testing.assert_array_equal(res["iterations"][cols], ref["iterations"], err_msg=tag)
    ok = ref["converged"]
    Vg = res["V"][:, cols][:, ok]
    dvm = np.abs(np.abs(Vg) - np.abs(ref["V"][:, ok])).max(initial=0.0)
    dva = np.abs(np.angle(Vg) - np.angle(ref["V"][:, ok])).max(initial=0.0)
    assert dvm <= TOL_V and dva <= TOL_V, (tag, dvm, dva)


def _bucket(n):
    return 0 if n <= 500 else (1 if n <= 5000 else 2)


def test_multi48_every_request_equals_alone(gpu):
    rng = np.random.default_rng(4)
    fam = grids.grid_family(48, 100, 20000, seed=4)
    counts = _counts(rng, 48, 4096, 1, 256)
    jobs = []
    for k, (g, b) in enumerate(zip(fam, counts)):
        Yp, Yj, Yx = g.csr()
        jobs.append((g, (Yp, Yj, Yx, g.pv, g.pq),
                     np.ascontiguousarray(grids.timeseries_scaled(g, int(b), seed=100 + k, sigma=0.1)),
                     np.ascontiguousarray(dc_initial_voltage(g))))
    srv = Server(max_slots=256)
    try:
        t0 = time.perf_counter()
        reqs = [srv.submit(*topo, S, V0) for _, topo, S, V0 in jobs]
        srv.wait_all(reqs, timeout=300)
        wall = time.perf_counter() - t0
    finally:
        srv.close()
    conv = sum(int(r.converged.sum()) for r in reqs)
    print(f"multi48: {int(counts.sum())} cases on {len(fam)} grids ({min(g.n for g in fam)}.."
          f"{max(g.n for g in fam)} buses) in {1e3 * wall:.0f} ms, {conv} converged")
    checked = [0, 0, 0]
    for k, ((g, topo, S, V0), r) in enumerate(zip(jobs, reqs)):
        res = r.result()
        alone = nr_klu.Solver(*topo).solve_batch(S, V0, max_iter=20, tol=1e-8)
        _same(res, alone, f"grid {k} ({g.n} buses, {S.shape[1]} cases)")
        bk = _bucket(g.n)
        if checked[bk] < 2:  # two grids per size bucket against the oracle
            cols = np.unique(np.linspace(0, S.shape[1] - 1, min(8, S.shape[1])).round().astype(int))
            _oracle_check(g, S, V0, res, cols, f"grid {k} ({g.n} buses)")
            checked[bk] += 1
    assert min(checked) >= 1, checked


def test_poisson_mix_up_to_25k_buses_equals_alone(gpu):
    rng = np.random.default_rng(5)
    fam = grids.grid_family(7, 100, 20000, seed=5) + [grids.sized_grid(25000, 3)]
    assert max(g.n for g in fam) >= 20000
    topos = []
    for g in fam:
        Yp, Yj, Yx = g.csr()
        topos.append((Yp, Yj, Yx, g.pv, g.pq))
    starts = [np.ascontiguousarray(dc_initial_voltage(g)) for g in fam]
    srv = Server(max_slots=512)
    nreq = 40
    arrivals = np.cumsum(rng.exponential(1.0 / 200.0, nreq))
    arrivals -= arrivals[0]
    jobs = []
    for k in range(nreq):
        gid = len(fam) - 1 if k % 8 == 0 else int(rng.integers(0, len(fam)))  # the 25k grid recurs
        g = fam[gid]
        b = int(np.clip(np.round(np.exp(rng.uniform(0.0, np.log(256)))), 1, 256))
        if rng.random() < 0.3:
            jobs.append(("n1", gid, np.sort(rng.choice(g.nbr, min(b, g.nbr), replace=False))))
        else:
            jobs.append(("ts", gid, np.ascontiguousarray(grids.timeseries_scaled(g, b, seed=1000 + k, sigma=0.1))))
    try:
        for gid, g in enumerate(fam):
            srv.register_outages(*topos[gid], g.f, g.t, g.branch_stamps(), contingency.single_branch_outages(g.nbr),
                                 g.ref, g.gen_z, True)
        reqs = []
        t0 = time.perf_counter()
        for k, (kind, gid, x) in enumerate(jobs):
            dt = t0 + arrivals[k] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            if kind == "ts":
                reqs.append(srv.submit(*topos[gid], x, starts[gid]))
            else:
                reqs.append(srv.submit_contingency(topology_key(*topos[gid]), x, fam[gid].sbus(), starts[gid]))
        srv.wait_all(reqs, timeout=300)
        st = srv.stats()
        print(f"poisson mix: {st['requests']} requests, {st['cases']} cases, p50 {st['p50_ms']:.0f} ms, "
              f"p99 {st['p99_ms']:.0f} ms")
        n1sets = {gid: srv.outage_sets[topology_key(*topos[gid])] for gid in range(len(fam))}
        oracle_done = False
        for (kind, gid, x), r in zip(jobs, reqs):
            res = r.result()
            g = fam[gid]
            if kind == "ts":
                alone = nr_klu.Solver(*topos[gid]).solve_batch(x, starts[gid], max_iter=20, tol=1e-8)
                _same(res, alone, f"ts grid {gid} ({g.n} buses)")
                if g.n >= 20000 and not oracle_done:
                    cols = np.unique(np.linspace(0, x.shape[1] - 1, min(4, x.shape[1])).round().astype(int))
                    _oracle_check(g, x, starts[gid], res, cols, f"{g.n}-bus grid")
                    oracle_done = True
            else:
                full = n1sets[gid]
                sub = contingency.ContingencySet(full.solver, full.f, full.t, *full.yval.T,
                                                 [full.groups[c] for c in x], full.ref, full.gen_z, full.reslack)
                alone = sub.solve(g.sbus(), starts[gid], max_iter=20, tol=1e-8)
                np.testing.assert_array_equal(res["served"], alone.served)
                _same(res, {"V": alone.V, "iterations": alone.iterations, "converged": alone.converged},
                      f"n1 grid {gid} ({g.n} buses)")
        assert oracle_done
    finally:
        srv.close()
