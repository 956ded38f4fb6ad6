"""ctypes wrapper of the CPU oracle (libnr_oracle.so) with the reference's nr_klu surface.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg. The product never imports this module.

``OracleSolver`` mirrors ``nr_klu.Solver`` (p3s/cpp/nr_klu.cpp:964-1313): the same
constructor arguments, ``update_Y``, ``lean_active``, ``solve``, ``solve_batch`` and
``solve_batch_contingency`` with the same shapes, dtypes, defaults and error strings.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libnr_oracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile libnr_oracle.so in place (gcc + OpenMP)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    L.oracle_topo_create.restype = C.c_void_p
    L.oracle_topo_create.argtypes = [_i32p, _i32p, _f64p, C.c_int, C.c_int, _i32p, C.c_int, _i32p, C.c_int,
                                     C.c_int, C.c_char_p, C.c_int]
    L.oracle_topo_free.argtypes = [C.c_void_p]
    L.oracle_lean_valid.argtypes = [C.c_void_p]
    L.oracle_dims.argtypes = [C.c_void_p, _i32p]
    L.oracle_rcond0.argtypes = [C.c_void_p]
    L.oracle_rcond0.restype = C.c_double
    L.oracle_update_Y.argtypes = [C.c_void_p, _f64p, C.c_int, C.c_char_p, C.c_int]
    L.oracle_solve_batch.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                     C.c_int, _f64p, _i32p, _u8p, C.c_char_p, C.c_int]
    L.oracle_solve_batch_contingency.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p, _u8p, C.c_int, C.c_int,
                                                 C.c_double, C.c_int, C.c_int, _f64p, _i32p, _u8p, C.c_char_p,
                                                 C.c_int]
    L.oracle_debug_J.argtypes = [C.c_void_p, _f64p, _f64p, _i32p, _i32p, _f64p, _f64p]
    L.oracle_gp_solve.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_int, _f64p]
    L.oracle_min_degree.argtypes = [C.c_int, _i32p, _i32p, _i32p]
    L.oracle_openmp.restype = C.c_int
    L.oracle_lean_rejections.restype = C.c_longlong
    L.oracle_column_reruns.restype = C.c_longlong
    L.oracle_backtracks.restype = C.c_longlong
    L.oracle_pivoting_factors.restype = C.c_longlong
    L.oracle_offdiagonal_pivots.restype = C.c_longlong
    L.oracle_set_lean_guard.argtypes = [C.c_double]
    _lib = L
    return L


def _c128_as_f64(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64).reshape(-1) if a.size else np.zeros(2)


class OracleSolver:
    """FP64 CPU restatement of ``nr_klu.Solver`` (KLU-free)."""

    def __init__(self, Yp, Yj, Yx, pv, pq, ordering: int = 0, btf: int = 0):
        L = lib()
        self.Yp = np.ascontiguousarray(Yp, dtype=np.int32)
        self.Yj = np.ascontiguousarray(Yj, dtype=np.int32)
        Yx = np.ascontiguousarray(Yx, dtype=np.complex128)
        self.pv = np.ascontiguousarray(pv, dtype=np.int32).ravel()
        self.pq = np.ascontiguousarray(pq, dtype=np.int32).ravel()
        self.n = int(self.Yp.shape[0] - 1)
        self.nnz = int(self.Yj.shape[0])
        err = C.create_string_buffer(256)
        self._h = L.oracle_topo_create(self.Yp, self.Yj, _c128_as_f64(Yx), self.n, self.nnz,
                                       self.pv if self.pv.size else np.zeros(1, np.int32), int(self.pv.size),
                                       self.pq if self.pq.size else np.zeros(1, np.int32), int(self.pq.size),
                                       int(ordering), err, 256)
        if not self._h:
            raise RuntimeError(err.value.decode())
        dims = np.zeros(5, dtype=np.int32)
        L.oracle_dims(self._h, dims)
        self.m, self.nnzJ, self.nnz_lu = int(dims[1]), int(dims[3]), int(dims[4])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.oracle_topo_free(h)
            self._h = None

    @property
    def lean_active(self) -> bool:
        return bool(lib().oracle_lean_valid(self._h))

    @property
    def rcond0(self) -> float:
        return float(lib().oracle_rcond0(self._h))

    def update_Y(self, Yx):
        Yx = np.ascontiguousarray(Yx, dtype=np.complex128).ravel()
        err = C.create_string_buffer(256)
        if lib().oracle_update_Y(self._h, _c128_as_f64(Yx), int(Yx.shape[0]), err, 256):
            raise RuntimeError(err.value.decode())

    def solve(self, Sbus, V0, max_iter: int = 30, tol: float = 1e-8, line_search: bool = True):
        S = np.ascontiguousarray(Sbus, dtype=np.complex128).reshape(self.n, 1)
        r = self.solve_batch(S, np.ascontiguousarray(V0, dtype=np.complex128), max_iter, tol, 1, line_search)
        return {"V": r["V"][:, 0].copy(), "iterations": int(r["iterations"][0]), "converged": bool(r["converged"][0])}

    def solve_batch(self, Sbus, V0, max_iter: int = 30, tol: float = 1e-8, n_threads: int = 0,
                    line_search: bool = True):
        S = np.ascontiguousarray(Sbus, dtype=np.complex128)
        if S.ndim != 2:
            raise RuntimeError("solve_batch: Sbus must be 2-D (n, T)")
        if S.shape[0] != self.n:
            raise RuntimeError("solve_batch: Sbus rows must equal n_bus")
        T = int(S.shape[1])
        V0 = np.ascontiguousarray(V0, dtype=np.complex128)
        per_col = V0.ndim == 2
        if per_col and V0.shape != (self.n, T):
            raise RuntimeError("solve_batch: V0 (n,T) shape mismatch")
        if not per_col and V0.shape[0] != self.n:
            raise RuntimeError("solve_batch: V0 (n,) length mismatch")
        V = np.zeros((self.n, max(T, 1)), dtype=np.complex128)
        it = np.zeros(max(T, 1), dtype=np.int32)
        cv = np.zeros(max(T, 1), dtype=np.uint8)
        err = C.create_string_buffer(256)
        rc = lib().oracle_solve_batch(self._h, _c128_as_f64(S), _c128_as_f64(V0), int(per_col), T, int(max_iter),
                                      float(tol), int(n_threads), int(bool(line_search)), V.view(np.float64).reshape(-1),
                                      it, cv, err, 256)
        if rc:
            raise RuntimeError("solve_batch: " + err.value.decode())
        return {"V": V[:, :T], "iterations": it[:T], "converged": cv[:T].astype(bool)}

    def solve_batch_contingency(self, Yx_mag, Yx_ang, Sbus, V0, pin, max_iter: int = 30, tol: float = 1e-8,
                                n_threads: int = 0, line_search: bool = True):
        Ym = np.ascontiguousarray(Yx_mag, dtype=np.float64)
        Ya = np.ascontiguousarray(Yx_ang, dtype=np.float64)
        if Ym.ndim != 2 or Ya.ndim != 2:
            raise RuntimeError("solve_batch_contingency: Yx_mag/ang must be 2-D (nnz, L)")
        if Ym.shape[0] != self.nnz:
            raise RuntimeError("solve_batch_contingency: Yx rows must equal Ybus nnz")
        L = int(Ym.shape[1])
        if Ya.shape != Ym.shape:
            raise RuntimeError("solve_batch_contingency: Yx_mag/ang shape mismatch")
        V0 = np.ascontiguousarray(V0, dtype=np.complex128)
        pin = np.ascontiguousarray(pin, dtype=np.uint8)
        if V0.ndim != 2 or V0.shape != (self.n, L):
            raise RuntimeError("solve_batch_contingency: V0 must be (n, L)")
        if pin.ndim != 2 or pin.shape != (self.n, L):
            raise RuntimeError("solve_batch_contingency: pin must be (n, L)")
        S = np.ascontiguousarray(Sbus, dtype=np.complex128).ravel()
        V = np.zeros((self.n, L), dtype=np.complex128)
        it = np.zeros(L, dtype=np.int32)
        cv = np.zeros(L, dtype=np.uint8)
        err = C.create_string_buffer(256)
        rc = lib().oracle_solve_batch_contingency(self._h, Ym.reshape(-1), Ya.reshape(-1), _c128_as_f64(S),
                                                  _c128_as_f64(V0), pin.reshape(-1), L, int(max_iter), float(tol),
                                                  int(n_threads), int(bool(line_search)),
                                                  V.view(np.float64).reshape(-1), it, cv, err, 256)
        if rc:
            raise RuntimeError("solve_batch_contingency: " + err.value.decode())
        return {"V": V, "iterations": it, "converged": cv.astype(bool)}

    def debug_J_at(self, Sbus, V0):
        Jp = np.zeros(self.m + 1, dtype=np.int32)
        Ji = np.zeros(self.nnzJ, dtype=np.int32)
        Jx = np.zeros(self.nnzJ, dtype=np.float64)
        F = np.zeros(self.m, dtype=np.float64)
        lib().oracle_debug_J(self._h, _c128_as_f64(Sbus), _c128_as_f64(V0), Jp, Ji, Jx, F)
        return {"Jp": Jp, "Ji": Ji, "Jx": Jx, "m": self.m, "F": F}


def debug_J(Yp, Yj, Yx, Sbus, V0, pv, pq):
    """The reference's module-level ``debug_J`` (p3s/cpp/nr_klu.cpp:1332-1342)."""
    return OracleSolver(Yp, Yj, Yx, pv, pq).debug_J_at(Sbus, V0)


def solve_single(Yp, Yj, Yx, Sbus, V0, pv, pq, max_iter=30, tol=1e-8, ordering=0, btf=0, line_search=True):
    """The reference's ``solve_single`` (p3s/cpp/nr_klu.cpp:868-951), without timings."""
    return OracleSolver(Yp, Yj, Yx, pv, pq, ordering=ordering).solve(Sbus, V0, max_iter, tol, line_search)


def gp_solve(A_csc, b, ordering: int = 0) -> np.ndarray:
    """Solve A x = b with the oracle's pivoting LU (stand-in for klu_factor + klu_solve)."""
    A = A_csc.tocsc()
    A.sort_indices()
    x = np.ascontiguousarray(b, dtype=np.float64).copy()
    rc = lib().oracle_gp_solve(A.shape[0], np.ascontiguousarray(A.indptr, np.int32),
                               np.ascontiguousarray(A.indices, np.int32), np.ascontiguousarray(A.data, np.float64),
                               int(ordering), x)
    if rc:
        raise np.linalg.LinAlgError("singular")
    return x


def min_degree(A_csc) -> np.ndarray:
    A = A_csc.tocsc()
    perm = np.zeros(A.shape[0], dtype=np.int32)
    lib().oracle_min_degree(A.shape[0], np.ascontiguousarray(A.indptr, np.int32),
                            np.ascontiguousarray(A.indices, np.int32), perm)
    return perm


def counters() -> dict:
    """Frozen-plan rejections, column re-runs, backtracking steps, fresh pivoting factorizations
    and the pivots those took off the diagonal, since the last reset_counters()."""
    L = lib()
    return {"lean_rejections": int(L.oracle_lean_rejections()), "column_reruns": int(L.oracle_column_reruns()),
            "backtracks": int(L.oracle_backtracks()), "pivoting_factors": int(L.oracle_pivoting_factors()),
            "offdiagonal_pivots": int(L.oracle_offdiagonal_pivots())}


def set_lean_guard(ratio: float = 1e-3) -> None:
    """Pivot-ratio bound below which the frozen plan is rejected (default 1e-3)."""
    lib().oracle_set_lean_guard(float(ratio))


def reset_counters() -> None:
    lib().oracle_reset_counters()



def _openmp() -> bool:
    return bool(lib().oracle_openmp())
