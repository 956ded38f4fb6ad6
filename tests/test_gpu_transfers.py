"""Large host transfers: above 16 MB the one-call entry points and resident batches move S,
V0, the per-case admittances and V through pinned slots that host threads pack and scatter
(slices of bus rows; column blocks when a call is split into chunks). The bytes must land
where the direct copies put them: V equals the calls small enough to copy directly, the
resident batch and the split call bit for bit, and sampled cases equal the oracle."""

import ctypes as C

import numpy as np
import pytest

from oracle.oracle import OracleSolver
from p3smi import _lib, grids, nr_klu
from p3smi.timeseries import dc_initial_voltage

pytestmark = pytest.mark.gpu

T = 4096  # ts2000: S, V0 and V are 131 MB each, 64 KB per bus row


def _same_bits(a, b, tag):
    assert np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64)), tag


def _near_oracle(r, ref, cols, tag):
    assert np.array_equal(r["converged"][cols], ref["converged"]), tag
    assert np.array_equal(r["iterations"][cols], ref["iterations"]), tag
    ok = ref["converged"]
    V = r["V"][:, cols][:, ok]
    assert np.abs(np.abs(V) - np.abs(ref["V"][:, ok])).max(initial=0.0) <= 1e-9, tag
    assert np.abs(np.angle(V) - np.angle(ref["V"][:, ok])).max(initial=0.0) <= 1e-9, tag


@pytest.fixture(scope="module")
def ts():
    g = grids.config_grid("ts2000", 0)
    Yp, Yj, Yx = g.csr()
    S = np.ascontiguousarray(grids.timeseries_daily(g, T, seed=3))
    rng = np.random.default_rng(4)
    # per-case starts: the DC start turned by up to 0.01 rad per bus (from a flat start a
    # quarter of these heavy hours do not converge in 20 iterations, in the oracle too)
    V0 = np.ascontiguousarray(dc_initial_voltage(g)[:, None] * np.exp(1j * rng.uniform(-0.01, 0.01, (g.n, T))))
    return g, (Yp, Yj, Yx), S, V0


def test_large_solve_batch_whole_split_and_resident(gpu, ts):
    g, (Yp, Yj, Yx), S, V0 = ts
    solver = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    r = solver.solve_batch(S, V0, max_iter=20, tol=1e-8)
    assert r["converged"].mean() > 0.99
    # the first and last 256 cases in calls of 8 MB (direct copies)
    for cols in (slice(0, 256), slice(T - 256, T)):
        rs = solver.solve_batch(np.ascontiguousarray(S[:, cols]), np.ascontiguousarray(V0[:, cols]),
                                max_iter=20, tol=1e-8)
        _same_bits(rs["V"], r["V"][:, cols], f"direct copies {cols}")
        assert np.array_equal(rs["iterations"], r["iterations"][cols])
    # split into chunks of 1408, 1344 and 1344 cases (whole waves of 64): column blocks packed
    # and scattered through the slots
    solver.set_max_chunk(1472)
    rc = solver.solve_batch(S, V0, max_iter=20, tol=1e-8)
    solver.set_max_chunk(0)
    _same_bits(rc["V"], r["V"], "split call")
    assert np.array_equal(rc["iterations"], r["iterations"])
    # a resident batch: upload, solve, download
    L = _lib.lib()
    bh = C.c_void_p()
    _lib.check(L.p3smi_batch_create(solver._h, T, C.byref(bh)))
    try:
        _lib.check(L.p3smi_batch_upload(bh, _lib.ptr(S), _lib.ptr(V0), 1))
        _lib.check(L.p3smi_batch_solve(bh, 20, 1e-8, 1, 0))
        V = np.empty((g.n, T), np.complex128)
        it = np.empty(T, np.int32)
        cv = np.empty(T, np.uint8)
        mis = np.empty(T, np.float64)
        _lib.check(L.p3smi_batch_download(bh, _lib.ptr(V), _lib.ptr(it), _lib.ptr(cv), _lib.ptr(mis)))
    finally:
        L.p3smi_batch_free(bh)
    _same_bits(V, r["V"], "resident batch")
    assert np.array_equal(it, r["iterations"])
    # sampled cases in every slice of rows and across the chunk borders against the oracle
    cols = np.r_[0:4, 1404:1412, 2748:2756, T - 4:T, 7:T:211]
    ref = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch(np.ascontiguousarray(S[:, cols]),
                                                           np.ascontiguousarray(V0[:, cols]), max_iter=20, tol=1e-8)
    _near_oracle(r, ref, cols, "oracle")


def test_large_contingency_batch_whole_and_split(gpu, ts):
    g, (Yp, Yj, Yx), S, V0 = ts
    L = 2048  # per-case admittances (nnzY, L): two 8-byte arrays above 16 MB
    rng = np.random.default_rng(5)
    # every case the whole matrix scaled by its own factor (scaling entries independently
    # leaves shunts the hours cannot be solved with, in the oracle too)
    Yc = Yx[:, None] * rng.uniform(0.97, 1.03, L)[None, :]
    mag, ang = np.abs(Yc), np.angle(Yc)
    assert mag.nbytes > (16 << 20)
    Sb = g.sbus()
    V0c = np.ascontiguousarray(V0[:, :L])
    pin = np.zeros((g.n, L), np.uint8)
    solver = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    r = solver.solve_batch_contingency(mag, ang, Sb, V0c, pin, max_iter=20, tol=1e-8)
    assert r["converged"].mean() > 0.99
    solver.set_max_chunk(704)  # chunks of 704, 704 and 640 cases
    rc = solver.solve_batch_contingency(mag, ang, Sb, V0c, pin, max_iter=20, tol=1e-8)
    solver.set_max_chunk(0)
    _same_bits(rc["V"], r["V"], "split contingency call")
    assert np.array_equal(rc["iterations"], r["iterations"])
    cols = np.r_[0:3, 700:708, 1404:1412, L - 3:L, 5:L:173]
    ref = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).solve_batch_contingency(
        np.ascontiguousarray(mag[:, cols]), np.ascontiguousarray(ang[:, cols]), Sb, np.ascontiguousarray(V0c[:, cols]),
        np.ascontiguousarray(pin[:, cols]), 20, 1e-8)
    _near_oracle(r, ref, cols, "oracle contingency")
