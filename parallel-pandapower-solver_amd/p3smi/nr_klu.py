"""Drop-in for the reference's ``nr_klu`` module, solved on the MI355X.

Same names, arguments, defaults, shapes, dtypes and error strings as the pybind11 module
``nr_klu`` (p3s/cpp/nr_klu.cpp:1315-1376), which the reference's callers import as
``from p3s import nr_klu`` (p3s/NewtonPowerflowCpp.py:29-32,
p3s/contingency/solver_cpp.py:36-39)::

    s = Solver(Yp, Yj, Yx, pv, pq)              # topology set-up once
    r = s.solve_batch(Sbus, V0, max_iter, tol)  # {V (n,T), iterations (T,), converged (T,)}

Every batch call runs the whole Newton iteration of all cases on the GPU through
libp3smi.so. Results add ``mismatch``: each case's final infinity-norm mismatch,
evaluated in FP64 on the device. ``n_threads`` is accepted for signature compatibility
(the work is not spread over CPU threads). ``precision`` selects the linear solver:
"double" (FP64 static-pivot LU, the default); "mixed" (FP32 LU refined in FP64 to the FP64
step, per-case FP64 fallback: the FP64 iteration, counts included); "mixed_switch" (the
reference GPU solver's rule: FP32 LU without refinement while the mismatch is above 1e-3,
FP64 after -- an inexact Newton whose iteration counts may differ).
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import P3smiError, check, lib, ptr

openmp = False  # the batch is spread over the GPU, not over OpenMP threads

_PREC = {"double": 0, "mixed": 1, "mixed_switch": 2}


def _c128(a, shape_err: str | None = None) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


def _prec(precision) -> int:
    try:
        return _PREC[precision]
    except KeyError:
        raise ValueError(f"precision must be 'double', 'mixed' or 'mixed_switch', got {precision!r}") from None


class Solver:
    """Batched polar Newton solver for one grid topology (``nr_klu.Solver``)."""

    def __init__(self, Yp, Yj, Yx, pv, pq, ordering: int = 0, btf: int = 0, n_slots: int = 0):
        self.Yp = np.ascontiguousarray(Yp, dtype=np.int32).ravel()
        self.Yj = np.ascontiguousarray(Yj, dtype=np.int32).ravel()
        self.Yx = _c128(Yx).ravel()
        self.pv = np.ascontiguousarray(pv, dtype=np.int32).ravel()
        self.pq = np.ascontiguousarray(pq, dtype=np.int32).ravel()
        self.n = int(self.Yp.shape[0] - 1)
        self.nnz = int(self.Yj.shape[0])
        if self.Yx.shape[0] != self.nnz:
            raise RuntimeError("Solver: Yx and Yj lengths differ")
        if ordering not in (0, 2):
            raise RuntimeError("Solver: ordering must be 0 (fill-reducing) or 2 (natural)")
        h = C.c_void_p()
        check(lib().p3smi_topology_create(ptr(self.Yp), ptr(self.Yj), ptr(self.Yx), self.n, self.nnz, ptr(self.pv),
                                          int(self.pv.size), ptr(self.pq), int(self.pq.size), int(ordering),
                                          int(n_slots), C.byref(h)), "Solver: ")
        self._h = h
        self.info = self._info()
        self.last_stats: dict = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().p3smi_topology_free(h)
            self._h = None

    def _fill_stats(self):
        """Counters of the last one-call solve (the reference's ``last_stats``,
        p3s/cuda/nr_polar_solver.py:671-681): Newton and line-search passes, FP32 / FP64
        factorizations, guard trips; summed over the chunks of a split call."""
        bh = C.c_void_p()
        check(lib().p3smi_topology_last_batch(self._h, C.byref(bh)))
        st = _lib.Stats()
        check(lib().p3smi_batch_stats(bh, C.byref(st)))
        self.last_stats = {f: getattr(st, f) for f, _ in st._fields_ if not f.startswith("reserved")}

    def _info(self) -> dict:
        inf = _lib.Info()
        check(lib().p3smi_topology_info(self._h, C.byref(inf)))
        return {f: getattr(inf, f) for f, _ in inf._fields_}

    def set_lu_shape(self, g: int, ns: int, few_g: int = 1, few_ns: int = 16, few_max: int = 256, group: int = 0):
        """LU launch shape: g cases per workgroup of 64*ns threads; (few_g, few_ns) while at
        most few_max cases are still iterating. group 0 at g = 64 factors 2x2 blocks of paired
        pivots per task; group > 0 deals runs of a pivot's entries together (1: one entry
        per task). Every shape gives the same bits."""
        check(lib().p3smi_topology_set_lu_shape(self._h, int(g), int(ns), int(few_g), int(few_ns), int(few_max),
                                                int(group)))
        self.info = self._info()

    def set_eval_tile(self, rows: int):
        """Ybus rows per workgroup tile of the mismatch/Jacobian evaluation (results do not
        depend on it)."""
        check(lib().p3smi_topology_set_eval_tile(self._h, int(rows)))
        self.info = self._info()

    def set_max_chunk(self, cases: int = 0):
        """Solve at most ``cases`` cases at once (a positive multiple of 64; 0: as many as 60%
        of the free HBM holds); larger calls run as nearly equal chunks one after the other
        (the reference GPU solver's ``max_chunk``, p3s/cuda/nr_polar_solver.py:311, 619-637)."""
        check(lib().p3smi_topology_set_max_chunk(self._h, int(cases)))

    def set_graphs(self, on: bool = True):
        """Replay each FP64 Newton iteration of later solves as a captured hipGraph."""
        check(lib().p3smi_topology_set_graphs(self._h, int(bool(on))))

    @property
    def lean_active(self) -> bool:
        """True when solves factor on the frozen static-pivot plan (p3s/cpp/nr_klu.cpp:1308): its
        flat-start factorization had no zero or non-finite pivot. Otherwise every case factors
        with pivoting."""
        return bool(self.info["rcond0"] > 0.0)

    def set_guard(self, ratio: float = 1e-3):
        """Pivot-ratio guard of the frozen plan (REUSE_RCOND_RATIO, p3s/cpp/nr_klu.cpp:564-566):
        a factorization whose pivot ratio falls below ratio x the flat-start one switches the
        case to the pivoting factorization; 1e30 rejects every frozen factorization."""
        check(lib().p3smi_topology_set_guard(self._h, float(ratio)))

    def set_pivoting(self, diagonal_first: bool = True):
        """How a fresh pivoting factorization starts: with diagonal_first (default) the
        static-pivot LU on the plan's order stands when the threshold rule would pick the
        diagonal at every column (the case then stays on the plan's order); otherwise, and
        always with False, the device Gilbert-Peierls LU chooses the pivots."""
        check(lib().p3smi_topology_set_pivoting(self._h, 1 if diagonal_first else 0))

    @property
    def m(self) -> int:
        return int(self.info["m"])

    def update_Y(self, Yx):
        """Refresh the Ybus values (same pattern) without redoing the symbolic set-up."""
        Yx = _c128(Yx).ravel()
        check(lib().p3smi_update_y(self._h, ptr(Yx), int(Yx.shape[0])))
        self.Yx = Yx
        self.info = self._info()

    def dc_start(self, Bdc, P, V_flat, ref) -> np.ndarray:
        """DC-power-flow start voltage computed on the device: B'(pvpq, pvpq) theta = P(pvpq) -
        B'(pvpq, ref) theta_ref solved with the engine's LU, |V_flat| kept at pv and pq buses
        (p3s/timeseries.py:200-244). Bdc: the DC susceptance matrix (n, n), its nonzeros within
        the Ybus pattern."""
        B = sp.csr_matrix(Bdc, dtype=np.float64)
        B.sum_duplicates()
        B.sort_indices()
        if B.shape != (self.n, self.n):
            raise RuntimeError("dc_start: Bdc must be (n, n)")
        n = np.int64(self.n)
        ykeys = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.Yp)) * n + self.Yj
        bkeys = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(B.indptr)) * n + B.indices
        pos = np.searchsorted(ykeys, bkeys)
        if bkeys.size and ((pos >= ykeys.size).any() or (ykeys[np.minimum(pos, ykeys.size - 1)] != bkeys).any()):
            raise RuntimeError("dc_start: Bdc has a nonzero outside the Ybus pattern")
        vals = np.zeros(self.nnz, dtype=np.float64)
        vals[pos] = B.data
        Pv = np.ascontiguousarray(P, dtype=np.float64).ravel()
        Vf = _c128(V_flat).ravel()
        ref = np.ascontiguousarray(ref, dtype=np.int32).ravel()
        if Pv.shape[0] != self.n or Vf.shape[0] != self.n:
            raise RuntimeError("dc_start: P and V_flat must be (n,)")
        out = np.empty(self.n, dtype=np.complex128)
        check(lib().p3smi_dc_start(self._h, ptr(vals), ptr(Pv), ptr(Vf), int(ref.size), ptr(ref), ptr(out)),
              "dc_start: ")
        return out

    # ---------------------------------------------------------------- solves
    def solve(self, Sbus, V0, max_iter: int = 30, tol: float = 1e-8, line_search: bool = True,
              precision: str = "double"):
        """One operating point: {V (n,), iterations int, converged bool}."""
        S = _c128(Sbus).reshape(self.n, 1)
        r = self.solve_batch(S, _c128(V0).ravel(), max_iter, tol, 0, line_search, precision)
        return {"V": r["V"][:, 0].copy(), "iterations": int(r["iterations"][0]),
                "converged": bool(r["converged"][0]), "mismatch": float(r["mismatch"][0])}

    def solve_batch(self, Sbus, V0, max_iter: int = 30, tol: float = 1e-8, n_threads: int = 0,
                    line_search: bool = True, precision: str = "double"):
        """Batch of operating points sharing this topology (p3s/cpp/nr_klu.cpp:1049-1166)."""
        S = _c128(Sbus)
        if S.ndim != 2:
            raise RuntimeError("solve_batch: Sbus must be 2-D (n, T)")
        if S.shape[0] != self.n:
            raise RuntimeError("solve_batch: Sbus rows must equal n_bus")
        T = int(S.shape[1])
        if T == 0:
            raise RuntimeError("solve_batch: T (time steps) must be > 0")
        if max_iter <= 0:
            raise RuntimeError("solve_batch: max_iter must be > 0")
        if tol <= 0.0:
            raise RuntimeError("solve_batch: tol must be > 0")
        if n_threads < 0:
            raise RuntimeError("solve_batch: n_threads must be >= 0")
        V0 = _c128(V0)
        per_col = V0.ndim == 2
        if per_col:
            if V0.shape != (self.n, T):
                raise RuntimeError("solve_batch: V0 (n,T) shape mismatch")
        elif V0.shape[0] != self.n:
            raise RuntimeError("solve_batch: V0 (n,) length mismatch")
        V = np.empty((self.n, T), dtype=np.complex128)
        it = np.empty(T, dtype=np.int32)
        cv = np.empty(T, dtype=np.uint8)
        mis = np.empty(T, dtype=np.float64)
        check(lib().p3smi_solve_batch(self._h, ptr(S), ptr(V0), int(per_col), T, int(max_iter), float(tol),
                                      int(bool(line_search)), _prec(precision), ptr(V), ptr(it), ptr(cv), ptr(mis)))
        self._fill_stats()
        return {"V": V, "iterations": it, "converged": cv.astype(bool), "mismatch": mis}

    def _check_contingency(self, L, Sbus, V0, pin, who="solve_batch_contingency"):
        V0 = _c128(V0)
        pin = np.ascontiguousarray(pin, dtype=np.uint8)
        if V0.ndim != 2 or V0.shape != (self.n, L):
            raise RuntimeError(f"{who}: V0 must be (n, L)")
        if pin.ndim != 2 or pin.shape != (self.n, L):
            raise RuntimeError(f"{who}: pin must be (n, L)")
        S = _c128(Sbus).ravel()
        if S.shape[0] != self.n:
            raise RuntimeError(f"{who}: Sbus must be (n,)")
        return S, V0, pin

    def _outputs(self, L):
        return (np.empty((self.n, L), dtype=np.complex128), np.empty(L, dtype=np.int32), np.empty(L, dtype=np.uint8),
                np.empty(L, dtype=np.float64))

    def solve_batch_contingency(self, Yx_mag, Yx_ang, Sbus, V0, pin, max_iter: int = 30, tol: float = 1e-8,
                                n_threads: int = 0, line_search: bool = True, precision: str = "double"):
        """N-1 batch with per-case Ybus magnitude/angle (nnz, L) (p3s/cpp/nr_klu.cpp:1184-1281)."""
        Ym = np.ascontiguousarray(Yx_mag, dtype=np.float64)
        Ya = np.ascontiguousarray(Yx_ang, dtype=np.float64)
        if Ym.ndim != 2 or Ya.ndim != 2:
            raise RuntimeError("solve_batch_contingency: Yx_mag/ang must be 2-D (nnz, L)")
        if Ym.shape[0] != self.nnz:
            raise RuntimeError("solve_batch_contingency: Yx rows must equal Ybus nnz")
        L = int(Ym.shape[1])
        if Ya.shape != Ym.shape:
            raise RuntimeError("solve_batch_contingency: Yx_mag/ang shape mismatch")
        S, V0, pin = self._check_contingency(L, Sbus, V0, pin)
        V, it, cv, mis = self._outputs(L)
        check(lib().p3smi_solve_batch_contingency(self._h, ptr(Ym), ptr(Ya), ptr(S), ptr(V0), ptr(pin), L,
                                                  int(max_iter), float(tol), int(bool(line_search)), _prec(precision),
                                                  ptr(V), ptr(it), ptr(cv), ptr(mis)))
        self._fill_stats()
        return {"V": V, "iterations": it, "converged": cv.astype(bool), "mismatch": mis}

    def solve_batch_contingency_cx(self, Yx, Sbus, V0, pin, max_iter: int = 30, tol: float = 1e-8,
                                   line_search: bool = True, precision: str = "double"):
        """N-1 batch with complex per-case Ybus values (nnz, L)
        (p3s/cuda/nr_polar_solver.py:483-509)."""
        Yx = _c128(Yx)
        if Yx.ndim != 2 or Yx.shape[0] != self.nnz:
            raise RuntimeError("solve_batch_contingency: Yx must be (Ybus nnz, L)")
        return self.solve_batch_contingency(np.abs(Yx), np.angle(Yx), Sbus, V0, pin, max_iter, tol, 0, line_search,
                                            precision)

    def solve_batch_contingency_stamp(self, stamp, Sbus, V0, pin, max_iter: int = 30, tol: float = 1e-8,
                                      line_search: bool = True, precision: str = "double"):
        """N-1 batch where case c runs on ``Yx - stamp[:, c]``; the per-case Ybus is expanded
        on the device (p3s/cuda/nr_polar_solver.py:531-557)."""
        st = sp.coo_matrix(stamp)
        if st.shape[0] != self.nnz:
            raise RuntimeError("solve_batch_contingency: stamp must be (Ybus nnz, L)")
        st.sum_duplicates()
        L = int(st.shape[1])
        S, V0, pin = self._check_contingency(L, Sbus, V0, pin)
        k = np.ascontiguousarray(st.row, dtype=np.int32)
        c = np.ascontiguousarray(st.col, dtype=np.int32)
        v = _c128(st.data)
        V, it, cv, mis = self._outputs(L)
        check(lib().p3smi_solve_batch_contingency_stamp(self._h, ptr(k), ptr(c), ptr(v), int(k.size), ptr(S), ptr(V0),
                                                        ptr(pin), L, int(max_iter), float(tol),
                                                        int(bool(line_search)), _prec(precision), ptr(V), ptr(it),
                                                        ptr(cv), ptr(mis)))
        self._fill_stats()
        return {"V": V, "iterations": it, "converged": cv.astype(bool), "mismatch": mis}

    # ---------------------------------------------------------------- host checks
    def jacobian_rows_at(self, Sbus, V0):
        """Jacobian in this engine's own layout (CSR by equation row, columns ascending) and
        the mismatch F = Scalc - Sspec at V0, evaluated on the host."""
        m, nnzJ = self.info["m"], self.info["nnz_j"]
        Jp = np.empty(m + 1, np.int32)
        Jj = np.empty(nnzJ, np.int32)
        Jx = np.empty(nnzJ, np.float64)
        R = np.empty(m, np.float64)
        check(lib().p3smi_host_jacobian(self._h, ptr(_c128(Sbus).ravel()), ptr(_c128(V0).ravel()), ptr(Jp), ptr(Jj),
                                        ptr(Jx), ptr(R)))
        return {"Jp": Jp, "Jj": Jj, "Jx": Jx, "m": m, "F": -R}

    def debug_J_at(self, Sbus, V0):
        """The Jacobian at V0 as the reference's ``debug_J_at`` returns it
        (p3s/cpp/nr_klu.cpp:1284-1304): CSC ``{Jp (m+1), Ji (nnz), Jx (nnz), m}`` in the order
        of ``build_J_pattern`` (p3s/cpp/nr_klu.cpp:230-285). Column c < npvpq is dVa of bus
        pvpq[c], the others dVm of bus pq[c - npvpq]; row r < npvpq is P of pvpq[r], the others
        Q of pq[r - npvpq]; within a column the P rows come first, then the Q rows, each in
        ascending bus order (the Ybus column order). ``F`` = Scalc - Sspec is added."""
        r = self.jacobian_rows_at(Sbus, V0)
        m = r["m"]
        npvpq = int(self.pv.size + self.pq.size)
        bus_of_row = np.concatenate([self.pv, self.pq, self.pq]).astype(np.int64)  # pvpq = [pv; pq]
        rows = np.repeat(np.arange(m, dtype=np.int64), np.diff(r["Jp"]))
        cols = r["Jj"].astype(np.int64)
        # sort entries by (column, Q after P, bus of the row)
        order = np.lexsort((bus_of_row[rows], rows >= npvpq, cols))
        Jp = np.zeros(m + 1, dtype=np.int32)
        Jp[1:] = np.cumsum(np.bincount(cols, minlength=m))
        return {"Jp": Jp, "Ji": np.ascontiguousarray(rows[order], dtype=np.int32),
                "Jx": np.ascontiguousarray(r["Jx"][order]), "m": m, "F": r["F"]}

    def host_factor_solve(self, Jx, R):
        """Factor J with the frozen plan and solve J x = R on the host (plan checks)."""
        x = np.empty(self.info["m"], np.float64)
        piv = np.empty(3, np.float64)
        check(lib().p3smi_host_factor_solve(self._h, ptr(np.ascontiguousarray(Jx, np.float64)),
                                            ptr(np.ascontiguousarray(R, np.float64)), ptr(x), ptr(piv)))
        return x, piv

    def host_lu_emulate(self, Jx, R, blocks: bool, nslots: int = 16, single: bool = False):
        """Host emulation of the device LU on the one-entry or the 2x2-block schedule (schedule
        checks): returns x (Jacobian column order), (min, max) |pivot|, and the check counts
        (violations, pivot pairs, operand steps)."""
        x = np.empty(self.info["m"], np.float64)
        piv = np.empty(2, np.float64)
        checks = np.zeros(3, np.int64)
        check(lib().p3smi_host_lu_emulate(self._h, int(blocks), int(nslots), int(single),
                                          ptr(np.ascontiguousarray(Jx, np.float64)),
                                          ptr(np.ascontiguousarray(R, np.float64)), ptr(x), ptr(piv), ptr(checks)))
        return x, piv, checks

    def maps(self):
        """(row_p, row_q, perm): Jacobian row of each bus's P / Q equation and the pivot order."""
        row_p = np.empty(self.n, np.int32)
        row_q = np.empty(self.n, np.int32)
        perm = np.empty(self.info["m"], np.int32)
        check(lib().p3smi_host_maps(self._h, ptr(row_p), ptr(row_q), ptr(perm)))
        return row_p, row_q, perm


def solve_single(Yp, Yj, Yx, Sbus, V0, pv, pq, max_iter: int = 30, tol: float = 1e-8, ordering: int = 0,
                 btf: int = 0, line_search: bool = True):
    """Stateless single solve (p3s/cpp/nr_klu.cpp:868-951)."""
    import time

    t0 = time.perf_counter()
    s = Solver(Yp, Yj, Yx, pv, pq, ordering=ordering, btf=btf)
    t1 = time.perf_counter()
    r = s.solve(Sbus, V0, max_iter, tol, line_search)
    t2 = time.perf_counter()
    r.update({"t_setup_ms": 1e3 * (t1 - t0), "t_solve_ms": 1e3 * (t2 - t1), "t_assemble_ms": 0.0, "profiled": False})
    return r


def debug_J(Yp, Yj, Yx, Sbus, V0, pv, pq):
    """Jacobian at V0 as CSC ``{Jp, Ji, Jx, m}`` in the reference's row and column order
    (p3s/cpp/nr_klu.cpp:1332-1342)."""
    return Solver(Yp, Yj, Yx, pv, pq).debug_J_at(Sbus, V0)


__all__ = ["Solver", "solve_single", "debug_J", "openmp", "P3smiError"]
