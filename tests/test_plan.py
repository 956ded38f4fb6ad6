"""Host-side checks of libp3smi.so that need no GPU: the library loads and exports every
symbol include/p3smi.h declares, the Jacobian pattern and values agree with the oracle's,
the frozen static-pivot plan factors and solves like a dense/sparse reference solve, and
bad inputs are rejected with the reference's messages."""

import os
import re

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle.oracle import OracleSolver
from p3smi import _lib, grids, nr_klu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    with open(os.path.join(ROOT, "include", "p3smi.h")) as fh:
        decl = set(re.findall(r"\b(p3smi_\w+)\s*\(", fh.read()))
    assert decl, "no declarations parsed"
    assert decl == set(_lib.EXPORTS)
    for name in decl:
        assert hasattr(L, name), name


def _random_V(g, seed):
    rng = np.random.default_rng(seed)
    return g.flat_start() * (1 + 0.03 * rng.standard_normal(g.n)) * np.exp(1j * 0.1 * rng.standard_normal(g.n))


@pytest.mark.parametrize("name", ["pf200", "ts2000"])
def test_jacobian_equals_oracle(name):
    g = grids.config_grid(name, 0)
    Yp, Yj, Yx = g.csr()
    V = _random_V(g, 1)
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    d = s.jacobian_rows_at(g.sbus(), V)
    J = sp.csr_matrix((d["Jx"], d["Jj"], d["Jp"]), shape=(d["m"], d["m"]))
    o = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).debug_J_at(g.sbus(), V)
    Jo = sp.csc_matrix((o["Jx"], o["Ji"], o["Jp"]), shape=(o["m"], o["m"]))
    assert (J != 0).nnz == (Jo != 0).nnz
    diff = abs(J - Jo).max()
    assert diff <= 1e-12 * abs(Jo).max()
    np.testing.assert_allclose(d["F"], o["F"], rtol=0, atol=1e-12 * max(1.0, np.abs(o["F"]).max()))


@pytest.mark.parametrize("name", ["pf200", "ts2000"])
def test_debug_J_has_the_reference_csc_layout(name):
    """debug_J returns the reference's CSC dict {Jp, Ji, Jx, m} (p3s/cpp/nr_klu.cpp:1284-1304):
    the same column pointers and row indices, entry for entry, as the oracle's restatement of
    build_J_pattern (p3s/cpp/nr_klu.cpp:230-285), and the same values; a caller building
    csc_matrix((Jx, Ji, Jp)) gets the Jacobian (tests/cuda/test_lean_lu.py:90-100)."""
    g = grids.config_grid(name, 0)
    Yp, Yj, Yx = g.csr()
    V = _random_V(g, 2)
    d = nr_klu.debug_J(Yp, Yj, Yx, g.sbus(), V, g.pv, g.pq)
    assert set(d) >= {"Jp", "Ji", "Jx", "m"}
    o = OracleSolver(Yp, Yj, Yx, g.pv, g.pq).debug_J_at(g.sbus(), V)
    assert d["m"] == o["m"]
    np.testing.assert_array_equal(d["Jp"], o["Jp"])
    np.testing.assert_array_equal(d["Ji"], o["Ji"])
    np.testing.assert_allclose(d["Jx"], o["Jx"], rtol=0, atol=1e-12 * np.abs(o["Jx"]).max())
    J = sp.csc_matrix((d["Jx"], d["Ji"], d["Jp"]), shape=(d["m"], d["m"]))
    Jr = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq).jacobian_rows_at(g.sbus(), V)
    Jc = sp.csr_matrix((Jr["Jx"], Jr["Jj"], Jr["Jp"]), shape=J.shape)
    assert abs(J - Jc).max() == 0.0


@pytest.mark.parametrize("name", ["pf200", "ts2000"])
def test_static_pivot_plan_solves_like_spsolve(name):
    g = grids.config_grid(name, 0)
    Yp, Yj, Yx = g.csr()
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    for seed in range(3):
        V = _random_V(g, seed)
        d = s.jacobian_rows_at(g.sbus(), V)
        J = sp.csr_matrix((d["Jx"], d["Jj"], d["Jp"]), shape=(d["m"], d["m"])).tocsc()
        R = -d["F"]
        x, piv = s.host_factor_solve(d["Jx"], R)
        xr = spla.spsolve(J, R)
        assert np.abs(J @ x - R).max() <= 1e-10 * max(1.0, np.abs(R).max())
        assert np.abs(x - xr).max() <= 1e-9 * max(1.0, np.abs(xr).max())
        assert np.isfinite(piv).all() and piv[0] > 0


def test_plan_statistics_are_consistent():
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    info = s.info
    assert info["m"] == len(g.pv) + 2 * len(g.pq)
    o = OracleSolver(Yp, Yj, Yx, g.pv, g.pq)
    assert info["nnz_j"] == o.nnzJ
    # fill of a symmetric static-pivot LU: nnz(L+U) >= nnz(J), every level non-empty
    assert info["nnz_lu"] >= info["nnz_j"]
    assert 0 < info["n_levels"] <= info["m"]
    assert 0 < info["rcond0"] <= 1.0
    # same pivot ratio as the oracle's frozen plan on the intact grid at flat start
    assert abs(info["rcond0"] - o.rcond0) <= 1e-9 * o.rcond0
    row_p, row_q, perm = s.maps()
    assert sorted(perm.tolist()) == list(range(info["m"]))
    assert (row_p[g.ref] == -1).all() and (row_q[g.pv] == -1).all()


def test_natural_ordering_gives_more_fill():
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    md = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq, ordering=0).info
    nat = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq, ordering=2).info
    assert md["nnz_lu"] < nat["nnz_lu"]


def test_invalid_inputs_are_rejected():
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    bad = Yj.copy()
    bad[[0, 1]] = bad[[1, 0]]
    with pytest.raises(RuntimeError, match="sorted"):
        nr_klu.Solver(Yp, bad, Yx, g.pv, g.pq)
    with pytest.raises(RuntimeError, match="twice"):
        nr_klu.Solver(Yp, Yj, Yx, g.pv, np.r_[g.pq, g.pv[:1]])
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    with pytest.raises(RuntimeError, match="update_Y: nnz mismatch"):
        s.update_Y(Yx[:-1])
    with pytest.raises(RuntimeError, match="solve_batch: Sbus must be 2-D"):
        s.solve_batch(g.sbus(), g.flat_start())
    with pytest.raises(RuntimeError, match="solve_batch: V0 \\(n,T\\) shape mismatch"):
        s.solve_batch(np.zeros((g.n, 3), complex), np.ones((g.n, 2), complex))
    with pytest.raises(ValueError, match="precision"):
        s.solve_batch(np.zeros((g.n, 3), complex), g.flat_start(), precision="half")
    for bad_chunk in (-64, 32, 100):
        with pytest.raises(RuntimeError, match="set_max_chunk: cases must be 0 or a positive multiple of 64"):
            s.set_max_chunk(bad_chunk)
    s.set_max_chunk(128)
    s.set_max_chunk(0)


def test_update_y_refreshes_reference_pivot_ratio():
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    r0 = s.info["rcond0"]
    s.update_Y(Yx * 1.5)
    assert abs(s.info["rcond0"] - r0) <= 1e-12  # scaling Y leaves the pivot ratio unchanged
    s.update_Y(Yx)
    assert s.info["rcond0"] == r0


@pytest.mark.skipif(_lib.device_count() > 0, reason="checks the no-GPU error path")
def test_solve_without_device_fails_loudly():
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    with pytest.raises(RuntimeError, match="no HIP device"):
        s.solve_batch(np.repeat(g.sbus()[:, None], 2, axis=1), g.flat_start())


@pytest.mark.parametrize("name", ["pf200", "ts2000"])
@pytest.mark.parametrize("single", [False, True])
def test_block_schedule_emulation_equals_one_entry_schedule(name, single):
    """The 2x2-block schedule, run on the host with the kernels' arithmetic, writes every
    factor position once, reads nothing before it is written (lists of a sub-level run last
    to first), and gives the one-entry schedule's solution bit for bit, in FP64 and FP32
    factors; the FP64 solution solves J x = R like scipy's sparse LU."""
    g = grids.config_grid(name, 0)
    Yp, Yj, Yx = g.csr()
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    for seed, nslots in ((0, 16), (1, 1024)):
        V = _random_V(g, seed)
        d = s.jacobian_rows_at(g.sbus(), V)
        R = -d["F"]
        x1, p1, c1 = s.host_lu_emulate(d["Jx"], R, blocks=False, nslots=nslots, single=single)
        x2, p2, c2 = s.host_lu_emulate(d["Jx"], R, blocks=True, nslots=nslots, single=single)
        assert c1[0] == 0 and c2[0] == 0, (c1, c2)
        assert c2[1] > 0.8 * len(g.pq) and c2[2] < c1[2]  # most PQ twins pair up; fewer steps than pairs
        assert np.array_equal(x1.view(np.int64), x2.view(np.int64))
        assert np.array_equal(p1, p2)
        if not single:
            J = sp.csr_matrix((d["Jx"], d["Jj"], d["Jp"]), shape=(d["m"], d["m"])).tocsc()
            xr = spla.spsolve(J, R)
            assert np.abs(x2 - xr).max() <= 1e-9 * max(1.0, np.abs(xr).max())
        else:
            x64, _, _ = s.host_lu_emulate(d["Jx"], R, blocks=True, nslots=nslots, single=False)
            assert np.abs(x2 - x64).max() <= 1e-3 * max(1.0, np.abs(x64).max())  # FP32 factors, FP64 sums


def test_block_emulation_rejects_bad_arguments():
    g = grids.config_grid("pf200", 0)
    Yp, Yj, Yx = g.csr()
    s = nr_klu.Solver(Yp, Yj, Yx, g.pv, g.pq)
    d = s.jacobian_rows_at(g.sbus(), g.flat_start())
    with pytest.raises(RuntimeError, match="nslots"):
        s.host_lu_emulate(d["Jx"], -d["F"], blocks=True, nslots=0)


def test_group_rejects_bad_arguments_without_a_device():
    import ctypes as C

    from p3smi import _lib

    L = _lib.lib()
    h = C.c_void_p()
    assert L.p3smi_group_create(None, 0, C.byref(h)) != 0
    assert b"at least one batch" in L.p3smi_last_error()
    assert L.p3smi_group_solve(None, 20, 1e-8, 1, 0, 8) != 0
    assert b"null group" in L.p3smi_last_error()
