"""Continuous batching: requests of two topologies arrive while others are on the GPU, are
admitted into free slots, and each result is bit-identical to solving that request alone."""

import time

import numpy as np
import pytest

from p3smi import grids, nr_klu
from p3smi.serving import Server
from p3smi.timeseries import dc_initial_voltage

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone(gpu):
    g1 = grids.config_grid("pf200", 0)
    g2 = grids.config_grid("ts2000", 0)
    topo = {}
    for name, g in (("pf200", g1), ("ts2000", g2)):
        Yp, Yj, Yx = g.csr()
        topo[name] = (g, Yp, Yj, Yx)
    S1 = grids.timeseries_scaled(g1, 600, seed=4)
    S2 = grids.timeseries_daily(g2, 800, seed=4)
    V1, V2 = g1.flat_start(), dc_initial_voltage(g2)
    rng = np.random.default_rng(0)
    srv = Server(max_slots=256, mem_budget=4 << 30)
    reqs = []
    try:
        c1 = c2 = 0
        for wave in range(4):
            for _ in range(3):
                b = int(rng.integers(1, 90))
                g, Yp, Yj, Yx = topo["pf200"]
                reqs.append(("pf200", c1, b, srv.submit(Yp, Yj, Yx, g.pv, g.pq, S1[:, c1:c1 + b], V1)))
                c1 += b
                b = int(rng.integers(1, 120))
                g, Yp, Yj, Yx = topo["ts2000"]
                reqs.append(("ts2000", c2, b, srv.submit(Yp, Yj, Yx, g.pv, g.pq, S2[:, c2:c2 + b], V2)))
                c2 += b
            time.sleep(0.02)  # the next wave arrives while these iterate
        srv.wait_all([r[3] for r in reqs], timeout=120)
        st = srv.stats()
        assert st["requests"] == len(reqs)
    finally:
        srv.close()
    solvers = {k: nr_klu.Solver(v[1], v[2], v[3], v[0].pv, v[0].pq) for k, v in topo.items()}
    for name, c0, b, r in reqs:
        S = S1 if name == "pf200" else S2
        V0 = V1 if name == "pf200" else V2
        alone = solvers[name].solve_batch(S[:, c0:c0 + b], V0, max_iter=20, tol=1e-8)
        res = r.result()
        np.testing.assert_array_equal(res["converged"], alone["converged"])
        np.testing.assert_array_equal(res["iterations"], alone["iterations"])
        np.testing.assert_array_equal(res["V"], alone["V"])


def test_n1_requests_and_pool_growth(gpu):
    """N-1 subsets served beside time-series requests equal the same subsets solved alone;
    a pool too small for a later request is rebuilt with room for it."""
    from p3smi import contingency

    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    V0 = g.flat_start()
    S = grids.timeseries_scaled(g, 400, seed=6)
    srv = Server(max_slots=512, mem_budget=2 << 30)
    try:
        key = srv.register_outages(Yp, Yj, Yx, g.pv, g.pq, g.f, g.t, g.branch_stamps(),
                                   contingency.single_branch_outages(g.nbr), g.ref, g.gen_z, True)
        small = srv.submit(Yp, Yj, Yx, g.pv, g.pq, S[:, :10], V0)
        small.wait(60)
        cap0 = srv.pools[key].capacity
        subsets = [np.arange(0, 40), np.arange(100, 103), np.arange(150, g.nbr)]
        n1 = [srv.submit_contingency(key, c, g.sbus(), V0) for c in subsets]
        big = srv.submit(Yp, Yj, Yx, g.pv, g.pq, S[:, 10:400], V0)
        srv.wait_all(n1 + [big], timeout=120)
        cap1 = srv.pools[key].capacity
    finally:
        srv.close()
    assert cap0 == 64 and cap1 == 448, (cap0, cap1)
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    for req, cols in ((small, slice(0, 10)), (big, slice(10, 400))):
        alone = s.solve_batch(S[:, cols], V0, max_iter=20, tol=1e-8)
        np.testing.assert_array_equal(req.result()["V"], alone["V"])
        np.testing.assert_array_equal(req.result()["iterations"], alone["iterations"])
    for req, c in zip(n1, subsets):
        cs = contingency.from_grid(g, groups=[np.array([k], np.int32) for k in c], reslack=True)
        alone = cs.solve(g.sbus(), V0, max_iter=20, tol=1e-8)
        res = req.result()
        np.testing.assert_array_equal(res["served"], alone.served)
        np.testing.assert_array_equal(res["iterations"], alone.iterations)
        np.testing.assert_array_equal(res["V"].view(np.float64), alone.V.view(np.float64))


def test_slot_batch_stats_count_factorizations_in_switching_mode(gpu):
    """A slot batch (continuous batching) driven pass by pass in the switching mixed mode
    reports its FP32 and FP64 factorizations, counted on the device: one per active case and
    Newton pass, plus an FP64 one for every FP32 pivot failure."""
    import ctypes as C

    from p3smi import _lib

    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    solver = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    L = _lib.lib()
    k = 40
    S = np.ascontiguousarray(grids.timeseries_scaled(g, k, seed=9), dtype=np.complex128)
    V0 = np.ascontiguousarray(np.repeat(g.flat_start()[:, None], k, axis=1), dtype=np.complex128)
    bh = C.c_void_p()
    _lib.check(L.p3smi_batch_create(solver._h, 64, C.byref(bh)))
    try:
        _lib.check(L.p3smi_batch_slots_init(bh))
        slots = np.arange(k, dtype=np.int32)
        _lib.check(L.p3smi_batch_load(bh, k, _lib.ptr(slots), _lib.ptr(S), _lib.ptr(V0)))
        counts = np.zeros(8, dtype=np.int32)
        for _ in range(30):
            _lib.check(L.p3smi_batch_pass(bh, 20, 1e-8, 1, 2, _lib.ptr(counts)))
            if counts[0] == 0 and counts[3] == 0 and counts[4] == 0 and counts[7] == 0:
                break
        buf = np.zeros(64, dtype=np.int32)
        nfin = C.c_int32()
        _lib.check(L.p3smi_batch_finished(bh, _lib.ptr(buf), C.byref(nfin)))
        assert nfin.value == k
        V = np.empty((g.n, k), dtype=np.complex128)
        it = np.empty(k, dtype=np.int32)
        cv = np.empty(k, dtype=np.uint8)
        mis = np.empty(k, dtype=np.float64)
        fin = np.ascontiguousarray(buf[:k])
        _lib.check(L.p3smi_batch_fetch(bh, k, _lib.ptr(fin), _lib.ptr(V), _lib.ptr(it), _lib.ptr(cv), _lib.ptr(mis)))
        assert cv.all()
        st = _lib.Stats()
        _lib.check(L.p3smi_batch_stats(bh, C.byref(st)))
        assert st.single_factorizations > 0
        assert st.single_factorizations + st.double_factorizations >= int(it.sum())
    finally:
        L.p3smi_batch_free(bh)


def test_worker_failure_fails_its_requests_and_the_pool_is_rebuilt(gpu):
    """A pass that fails inside a pool's worker thread ends every request the pool holds with
    the error (no request waits forever); the next request of that topology gets a new pool."""
    from p3smi import _lib

    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    S = grids.timeseries_scaled(g, 40, seed=6)
    V0 = g.flat_start()
    srv = Server(max_slots=64, mem_budget=1 << 30)
    try:
        ok = srv.submit(Yp, Yj, Yx, g.pv, g.pq, S[:, :8], V0)
        srv.wait_all([ok], timeout=60)
        pool = next(iter(srv.pools.values()))

        class Failing:
            def __init__(self, L):
                self._L = L

            def __getattr__(self, name):
                if name == "p3smi_batch_pass":
                    return lambda *a: 1  # P3SMI_E_ARG
                return getattr(self._L, name)

        pool.L = Failing(pool.L)
        bad = [srv.submit(Yp, Yj, Yx, g.pv, g.pq, S[:, 8 + 4 * k:12 + 4 * k], V0) for k in range(3)]
        for r in bad:
            assert r._done.wait(60), "a request of the failed pool never finished"
            with pytest.raises(_lib.P3smiError):
                r.wait()
        pool.L = pool.L._L
        again = srv.submit(Yp, Yj, Yx, g.pv, g.pq, S[:, 20:40], V0)
        srv.wait_all([again], timeout=60)
        assert next(iter(srv.pools.values())) is not pool
        alone = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(S[:, 20:40], V0, max_iter=20, tol=1e-8)
        np.testing.assert_array_equal(again.result()["V"], alone["V"])
    finally:
        srv.close()
