"""Sbus built on the device (p3s/timeseries.py:36-141) equals the numpy restatement bit for
bit, and solving from it equals solving from the host-built matrix."""

import ctypes as C

import numpy as np
import pytest

from oracle.sbus import build_sbus
from p3smi import _lib, grids, nr_klu
from p3smi.timeseries import SbusBuilder, dc_initial_voltage

pytestmark = pytest.mark.gpu


def _sources(g, T, seed):
    """Loads (p, q) at every PQ bus and static generators (p) at every PV bus, in p.u. with
    the reference's signs (loads -1, generation +1, per unit of sn = 1), daily profiles."""
    rng = np.random.default_rng(seed)
    t = np.arange(T)
    prof = 1.0 + 0.25 * np.sin(2 * np.pi * (t % 24) / 24.0)
    lp = g.p_load[g.pq][:, None] * prof[None, :] * rng.normal(1.0, 0.05, (g.pq.size, T))
    lq = g.q_load[g.pq][:, None] * prof[None, :] * rng.normal(1.0, 0.05, (g.pq.size, T))
    gp = g.p_gen[g.pv][:, None] * rng.uniform(0.0, 1.0, (g.pv.size, T))
    scaling = rng.uniform(0.9, 1.1, g.pq.size)
    sources = [
        {"bus": g.pq, "coef": -1.0 * scaling, "base": g.p_load[g.pq], "imag": False},
        {"bus": g.pq, "coef": -1.0 * scaling, "base": g.q_load[g.pq], "imag": True},
        {"bus": g.pv, "coef": np.ones(g.pv.size), "base": g.p_gen[g.pv], "imag": False},
    ]
    return sources, [lp, lq, gp]


def test_device_sbus_equals_restatement_and_solves_alike(gpu):
    g = grids.config_grid("ts2000", 0)
    Yp, Yj, Yx = g.csr()
    T = 640
    s0 = g.sbus()
    sources, series = _sources(g, T, 5)
    host = build_sbus(g.n, s0, sources, series)
    sb = SbusBuilder(g.n, s0, sources)
    solver = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    L = _lib.lib()
    V0 = np.ascontiguousarray(dc_initial_voltage(g))
    bh = C.c_void_p()
    _lib.check(L.p3smi_batch_create(solver._h, T, C.byref(bh)))
    try:
        sb.build(bh, series)
        S_dev = np.empty((g.n, T), dtype=np.complex128)
        _lib.check(L.p3smi_batch_download_sbus(bh, _lib.ptr(S_dev)))
        assert np.array_equal(S_dev.view(np.float64), host.view(np.float64))
        _lib.check(L.p3smi_batch_upload(bh, None, _lib.ptr(V0), 0))  # keep the device-built S
        _lib.check(L.p3smi_batch_solve(bh, 20, 1e-8, 1, 0))
        V = np.empty((g.n, T), dtype=np.complex128)
        it = np.empty(T, dtype=np.int32)
        cv = np.empty(T, dtype=np.uint8)
        mis = np.empty(T, dtype=np.float64)
        _lib.check(L.p3smi_batch_download(bh, _lib.ptr(V), _lib.ptr(it), _lib.ptr(cv), _lib.ptr(mis)))
        # the batch no longer holds a device-built S: a second NULL upload is refused
        assert L.p3smi_batch_upload(bh, None, _lib.ptr(V0), 0) != 0
    finally:
        L.p3smi_batch_free(bh)
    ref = solver.solve_batch(host, V0, max_iter=20, tol=1e-8)
    assert cv.astype(bool).all()
    assert np.array_equal(it, ref["iterations"])
    assert np.array_equal(V.view(np.float64), ref["V"].view(np.float64))


def test_device_sbus_large_series_both_entry_points(gpu):
    """4096 cases: the load series (a row per PQ bus, above 16 MB each) go through the pinned
    slots. The block entry point (series one by one) and the stacked-array entry point write
    the same S as the restatement."""
    g = grids.config_grid("ts2000", 0)
    Yp, Yj, Yx = g.csr()
    T = 4096
    s0 = g.sbus()
    sources, series = _sources(g, T, 6)
    host = build_sbus(g.n, s0, sources, series)
    sb = SbusBuilder(g.n, s0, sources)
    solver = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    L = _lib.lib()
    bh = C.c_void_p()
    _lib.check(L.p3smi_batch_create(solver._h, T, C.byref(bh)))
    try:
        S_dev = np.empty((g.n, T), dtype=np.complex128)
        sb.build(bh, series)
        _lib.check(L.p3smi_batch_download_sbus(bh, _lib.ptr(S_dev)))
        assert np.array_equal(S_dev.view(np.float64), host.view(np.float64))
        X = sb.stack(series)
        _lib.check(L.p3smi_batch_build_sbus(bh, _lib.ptr(sb.s0), _lib.ptr(sb.ptr), int(sb.row.size), _lib.ptr(sb.row),
                                            _lib.ptr(sb.coef), _lib.ptr(sb.base), _lib.ptr(sb.imag), int(sb.nrows),
                                            _lib.ptr(X)))
        S_dev[...] = 0
        _lib.check(L.p3smi_batch_download_sbus(bh, _lib.ptr(S_dev)))
        assert np.array_equal(S_dev.view(np.float64), host.view(np.float64))
    finally:
        L.p3smi_batch_free(bh)


def test_device_sbus_blocks_rejects_bad_arguments(gpu):
    """The block entry point checks its blocks before any copy: a negative row count, a missing
    block pointer, and contributions indexing past the rows the blocks add up to."""
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    T = 8
    sources, series = _sources(g, T, 7)
    sb = SbusBuilder(g.n, g.sbus(), sources)
    X = sb.blocks(series)
    solver = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    L = _lib.lib()
    bh = C.c_void_p()
    _lib.check(L.p3smi_batch_create(solver._h, T, C.byref(bh)))

    def call(rows, ptrs, cols=T):
        rows = np.asarray(rows, dtype=np.int32)
        blk = (C.c_void_p * max(len(ptrs), 1))(*ptrs)
        return L.p3smi_batch_build_sbus_blocks(bh, _lib.ptr(sb.s0), _lib.ptr(sb.ptr), int(sb.row.size),
                                               _lib.ptr(sb.row), _lib.ptr(sb.coef), _lib.ptr(sb.base),
                                               _lib.ptr(sb.imag), int(cols), len(rows), _lib.ptr(rows),
                                               C.cast(blk, C.c_void_p))

    try:
        good_rows = [x.shape[0] for x in X]
        good_ptrs = [x.ctypes.data for x in X]
        assert call(good_rows, good_ptrs) == 0
        bad = list(good_rows)
        bad[0] = -1
        assert call(bad, good_ptrs) != 0
        assert "negative" in L.p3smi_last_error().decode()
        assert call(good_rows, [good_ptrs[0], None, good_ptrs[2]]) != 0
        assert "null" in L.p3smi_last_error().decode()
        assert call(good_rows[:2], good_ptrs[:2]) != 0  # the generator rows are missing
        assert "row index out of range" in L.p3smi_last_error().decode()
        # series narrower or wider than the batch: rejected before any copy
        for cols in (T - 1, T + 1):
            assert call(good_rows, good_ptrs, cols) != 0
            assert "columns" in L.p3smi_last_error().decode()
        narrow = [np.ascontiguousarray(x[:, :T - 2]) for x in X]
        with pytest.raises(Exception, match="columns"):
            sb.build(bh, narrow)
    finally:
        L.p3smi_batch_free(bh)
