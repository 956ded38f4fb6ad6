"""N-1 contingency analysis on the device: outage cases, islands, reference buses, solves.

The host side restates the input preparation of the reference's case generator
(p3s/contingency/case_generator.py) for branch lists given directly in p.u.:

  * every branch's 2x2 stamp and the four Ybus CSR positions it lands on
    (``_index_branches``, :270-317);
  * live branches: a branch couples its ends iff its Y_ft or Y_tf stamp is non-zero (:316);
  * one outage case per group of branches; ``single_branch_outages`` gives the N-1 set.

Everything per case then runs on the GPU (libp3smi.so): the island labels by union-find,
the reference of every island (lowest-index slack, else with re-slacking the lowest-|Z|
generator, ties within np.isclose to the lowest bus index), the served mask, the pin mask
and start voltage (p3s/contingency/solver_cpp.py:119-134), the per-case Ybus (base minus
the outaged stamps) and the Newton solve. Voltages come back NaN at unserved buses.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib, ptr
from .nr_klu import Solver


@dataclass
class ContingencyResultTable:
    """Per-case results (p3s/contingency/solver_cpp.py:47-61)."""

    groups: list
    V: np.ndarray  # (n, L) complex, NaN at unserved buses
    served: np.ndarray  # (n, L) bool
    converged: np.ndarray  # (L,) bool
    iterations: np.ndarray  # (L,) int
    mismatch: np.ndarray  # (L,) final infinity-norm mismatch (FP64, on the device)

    @property
    def vm(self) -> np.ndarray:
        return np.abs(self.V)

    @property
    def va(self) -> np.ndarray:
        return np.degrees(np.angle(self.V))


def branch_positions(Yp, Yj, f, t) -> np.ndarray:
    """(nbr, 4) Ybus CSR data positions of (f,f), (f,t), (t,f), (t,t) for every branch."""
    Yp = np.asarray(Yp, dtype=np.int64)
    Yj = np.asarray(Yj, dtype=np.int64)
    n = Yp.shape[0] - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(Yp))
    keys = rows * n + Yj  # ascending: rows ascending, columns sorted within a row
    f = np.asarray(f, dtype=np.int64)
    t = np.asarray(t, dtype=np.int64)
    want = np.stack([f * n + f, f * n + t, t * n + f, t * n + t], axis=1)
    pos = np.searchsorted(keys, want)
    if (pos >= keys.shape[0]).any() or (keys[np.minimum(pos, keys.shape[0] - 1)] != want).any():
        raise ValueError("a branch stamp entry is not a structural nonzero of Ybus")
    return pos.astype(np.int32)


def single_branch_outages(nbr: int) -> list:
    """The N-1 set: one case per branch."""
    return [np.array([k], dtype=np.int32) for k in range(nbr)]


class ContingencySet:
    """Outage cases of one topology, resident on the device."""

    def __init__(self, solver: Solver, f, t, yff, yft, ytf, ytt, groups, ref, gen_z=None, reslack: bool = False,
                 names=None):
        self.solver = solver
        self.n = solver.n
        self.f = np.ascontiguousarray(f, dtype=np.int32)
        self.t = np.ascontiguousarray(t, dtype=np.int32)
        self.nbr = int(self.f.shape[0])
        self.ypos = np.ascontiguousarray(branch_positions(solver.Yp, solver.Yj, self.f, self.t))
        self.yval = np.ascontiguousarray(np.stack([yff, yft, ytf, ytt], axis=1).astype(np.complex128))
        self.live = (self.yval[:, 1] != 0) | (self.yval[:, 2] != 0)
        self.groups = [np.asarray(g, dtype=np.int32).ravel() for g in groups]
        self.names = list(names) if names is not None else list(range(len(self.groups)))
        self.L = len(self.groups)
        gptr = np.zeros(self.L + 1, dtype=np.int32)
        gptr[1:] = np.cumsum([g.size for g in self.groups])
        gbr = np.concatenate(self.groups).astype(np.int32) if self.L else np.zeros(0, np.int32)
        self.ref = np.ascontiguousarray(ref, dtype=np.int32).ravel()
        gz = None if gen_z is None else np.ascontiguousarray(gen_z, dtype=np.float64)
        if gz is not None and gz.shape != (self.n,):
            raise ValueError("gen_z must be (n_bus,) with +inf where a bus has no eligible generator")
        self.gen_z = gz
        self.reslack = bool(reslack)
        h = C.c_void_p()
        check(lib().p3smi_outages_create(solver._h, self.nbr, ptr(self.f), ptr(self.t), ptr(self.ypos),
                                         ptr(self.yval), self.L, ptr(gptr), ptr(gbr if gbr.size else np.zeros(1, np.int32)),
                                         int(self.ref.size), ptr(self.ref), ptr(gz), int(self.reslack), C.byref(h)),
              "contingency: ")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().p3smi_outages_free(h)
            self._h = None

    def islands(self):
        """served (n, L) bool, ref_bus (n, L) int32 (-1 = none), re-slacked references (L,)."""
        served = np.empty((self.n, self.L), dtype=np.uint8)
        refb = np.empty((self.n, self.L), dtype=np.int32)
        pinned = np.empty(self.L, dtype=np.int32)
        check(lib().p3smi_outages_islands(self._h, ptr(served), ptr(refb), ptr(pinned)))
        return served.astype(bool), refb, pinned

    def case_ybus(self, c: int) -> np.ndarray:
        """Ybus values of case c (host, for checks): base minus the outaged stamps."""
        Yx = self.solver.Yx.copy()
        for k in self.groups[c]:
            for q in range(4):
                Yx[self.ypos[k, q]] -= self.yval[k, q]
        return Yx

    def solve(self, Sbus, V_start, max_iter: int = 30, tol: float = 1e-8, line_search: bool = True,
              precision: str = "double") -> ContingencyResultTable:
        from .nr_klu import _prec

        S = np.ascontiguousarray(Sbus, dtype=np.complex128).ravel()
        V0 = np.ascontiguousarray(V_start, dtype=np.complex128).ravel()
        if S.shape != (self.n,) or V0.shape != (self.n,):
            raise RuntimeError("solve_batch_contingency: Sbus and V_start must be (n_bus,)")
        V = np.empty((self.n, self.L), dtype=np.complex128)
        it = np.empty(self.L, dtype=np.int32)
        cv = np.empty(self.L, dtype=np.uint8)
        mis = np.empty(self.L, dtype=np.float64)
        served = np.empty((self.n, self.L), dtype=np.uint8)
        check(lib().p3smi_outages_solve(self._h, ptr(S), ptr(V0), int(max_iter), float(tol), int(bool(line_search)),
                                        _prec(precision), ptr(V), ptr(it), ptr(cv), ptr(mis), ptr(served)))
        return ContingencyResultTable(groups=self.names, V=V, served=served.astype(bool), converged=cv.astype(bool),
                                      iterations=it, mismatch=mis)


def from_grid(grid, groups=None, reslack: bool = False, solver: Solver | None = None) -> ContingencySet:
    """Outage set of a synthetic grid (default: every single-branch outage)."""
    if solver is None:
        Yp, Yj, Yx = grid.csr()
        solver = Solver(Yp, Yj, Yx, grid.pv, grid.pq)
    yff, yft, ytf, ytt = grid.branch_stamps()
    if groups is None:
        groups = single_branch_outages(grid.nbr)
    return ContingencySet(solver, grid.f, grid.t, yff, yft, ytf, ytt, groups, grid.ref, grid.gen_z, reslack)


def solve_contingencies(grid, reslack: bool = False, tol: float = 1e-8, max_iter: int = 30, init: str = "dc",
                        precision: str = "double", groups=None) -> ContingencyResultTable:
    """Every outage case of the grid on the device (p3s/contingency/solver_cpp.py:64-156)."""
    from .timeseries import dc_initial_voltage

    cs = from_grid(grid, groups=groups, reslack=reslack)
    if init == "dc":
        v_start = dc_initial_voltage(grid, solver=cs.solver)  # the DC solve on the device
    elif init == "flat":
        v_start = grid.flat_start()
    else:
        raise ValueError(f"init must be 'dc' or 'flat', got {init!r}")
    return cs.solve(grid.sbus(), v_start, max_iter=max_iter, tol=tol, precision=precision)


# ---------------------------------------------------------------------- post-processing
class _BranchTable(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in ("p_from", "q_from", "p_to", "q_to", "i_from", "i_to", "i_ka", "loading")]


@dataclass
class BranchLoadingTable:
    """(n_branch, L) branch results (p3s/contingency/find_overloads.py:48-100). NaN: the case
    did not converge or the branch touches an unserved bus; 0: the branch is the outage."""

    groups: list
    p_from_mw: np.ndarray
    q_from_mvar: np.ndarray
    p_to_mw: np.ndarray
    q_to_mvar: np.ndarray
    i_from_ka: np.ndarray
    i_to_ka: np.ndarray
    i_ka: np.ndarray
    loading_percent: np.ndarray
    max_loading_percent: np.ndarray  # (n_branch,) nanmax over the cases
    worst_case: np.ndarray  # (n_branch,) first case of the maximum, -1 if all NaN


@dataclass
class OverloadReport:
    """Every (branch, case) loaded above the threshold, worst first (ties: branch, case)."""

    threshold_percent: float
    branch: np.ndarray
    case: np.ndarray
    group: np.ndarray
    loading_percent: np.ndarray
    i_ka: np.ndarray
    rating: np.ndarray
    max_loading_percent: np.ndarray
    worst_case: np.ndarray

    def __len__(self) -> int:
        return int(self.branch.size)


@dataclass
class VoltageViolationReport:
    """Every (bus, case) outside its band, ordered as p3s/contingency/voltage_band.py:164-182."""

    bus: np.ndarray
    case: np.ndarray
    group: np.ndarray
    kind: np.ndarray  # "low" / "high"
    vm_pu: np.ndarray
    limit_pu: np.ndarray
    deviation_pu: np.ndarray

    def __len__(self) -> int:
        return int(self.bus.size)


def _per_bus(v, n, name):
    a = np.asarray(v, dtype=np.float64)
    if a.ndim == 0:
        return np.full(n, float(a))
    if a.shape != (n,):
        raise ValueError(f"{name} must be a scalar or have shape ({n},)")
    return np.ascontiguousarray(a)


def _branch_args(cs, vn_from_kv, vn_to_kv, rating, trafo):
    nb = cs.nbr
    vf = _per_bus(vn_from_kv, nb, "vn_from_kv")
    vt = vf if vn_to_kv is None else _per_bus(vn_to_kv, nb, "vn_to_kv")
    rt = _per_bus(rating, nb, "rating")
    tr = np.zeros(nb, np.uint8) if trafo is None else np.ascontiguousarray(np.asarray(trafo, dtype=bool), dtype=np.uint8)
    if tr.shape != (nb,):
        raise ValueError(f"trafo must have shape ({nb},)")
    return vf, vt, rt, tr


def _branch_flows(cs, vn_from_kv, vn_to_kv, rating, trafo, sn_mva, threshold, table, cap):
    vf, vt, rt, tr = _branch_args(cs, vn_from_kv, vn_to_kv, rating, trafo)
    best = np.empty(cs.nbr, np.float64)
    arg = np.empty(cs.nbr, np.int32)
    ob = np.empty(max(cap, 1), np.int32)
    oc = np.empty(max(cap, 1), np.int32)
    ol = np.empty(max(cap, 1), np.float64)
    oi = np.empty(max(cap, 1), np.float64)
    n = C.c_int32(0)
    check(lib().p3smi_outages_branch_flows(cs._h, ptr(vf), ptr(vt), ptr(rt), ptr(tr), float(sn_mva), float(threshold),
                                           C.byref(table) if table is not None else None, ptr(best), ptr(arg),
                                           int(cap), ptr(ob), ptr(oc), ptr(ol), ptr(oi), C.byref(n)),
          "contingency: ")
    return best, arg, int(n.value), ob, oc, ol, oi, rt


def branch_loading(cs, vn_from_kv, vn_to_kv=None, rating=1.0, trafo=None, sn_mva: float = 1.0) -> BranchLoadingTable:
    """compute_branch_loading (p3s/contingency/find_overloads.py:128-262) of the set's last
    solve, on the device; the (n_branch, L) tables come back to the host."""
    shape = (cs.nbr, cs.L)
    arrs = {f: np.empty(shape, np.float64) for f, _ in _BranchTable._fields_}
    tab = _BranchTable(*[arrs[f].ctypes.data for f, _ in _BranchTable._fields_])
    best, arg, *_ = _branch_flows(cs, vn_from_kv, vn_to_kv, rating, trafo, sn_mva, np.inf, tab, 0)
    return BranchLoadingTable(groups=cs.names, p_from_mw=arrs["p_from"], q_from_mvar=arrs["q_from"],
                              p_to_mw=arrs["p_to"], q_to_mvar=arrs["q_to"], i_from_ka=arrs["i_from"],
                              i_to_ka=arrs["i_to"], i_ka=arrs["i_ka"], loading_percent=arrs["loading"],
                              max_loading_percent=best, worst_case=arg.astype(np.int64))


def find_overloads(cs, threshold_percent: float, vn_from_kv, vn_to_kv=None, rating=1.0, trafo=None,
                   sn_mva: float = 1.0, cap: int = 1 << 20) -> OverloadReport:
    """find_overloads (p3s/contingency/find_overloads.py:322-420) of the set's last solve: only
    the violations and the per-branch worst case leave the device."""
    if not np.isfinite(threshold_percent):
        raise ValueError(f"threshold_percent must be finite, got {threshold_percent!r}")
    best, arg, n, ob, oc, ol, oi, rt = _branch_flows(cs, vn_from_kv, vn_to_kv, rating, trafo, sn_mva,
                                                     threshold_percent, None, cap)
    if n > cap:
        raise RuntimeError(f"find_overloads: {n} violations exceed cap={cap}; raise cap or the threshold")
    ob, oc, ol, oi = ob[:n], oc[:n], ol[:n], oi[:n]
    order = np.lexsort((oc, ob, -ol))  # worst first, ties by branch then case
    groups = np.asarray(cs.names, dtype=object)
    return OverloadReport(threshold_percent=float(threshold_percent), branch=ob[order].astype(np.int64),
                          case=oc[order].astype(np.int64), group=groups[oc[order]], loading_percent=ol[order],
                          i_ka=oi[order], rating=rt[ob[order]], max_loading_percent=best,
                          worst_case=arg.astype(np.int64))


def find_voltage_violations(cs, vm_min_pu, vm_max_pu, in_service=None, tol_pu: float = 1e-9,
                            cap: int = 1 << 20) -> VoltageViolationReport:
    """find_voltage_violations (p3s/contingency/voltage_band.py:103-182) of the set's last solve."""
    if not np.isfinite(tol_pu) or tol_pu < 0:
        raise ValueError(f"tol_pu must be finite and >= 0, got {tol_pu!r}")
    n_bus = cs.n
    lo = _per_bus(vm_min_pu, n_bus, "vm_min_pu")
    hi = _per_bus(vm_max_pu, n_bus, "vm_max_pu")
    if np.isnan(lo).all() and np.isnan(hi).all():
        raise ValueError("no voltage band: pass vm_min_pu / vm_max_pu")
    ins = np.ones(n_bus, np.uint8) if in_service is None else np.ascontiguousarray(in_service, dtype=np.uint8)
    ob = np.empty(max(cap, 1), np.int32)
    oc = np.empty(max(cap, 1), np.int32)
    ov = np.empty(max(cap, 1), np.float64)
    ok = np.empty(max(cap, 1), np.uint8)
    n = C.c_int32(0)
    check(lib().p3smi_outages_voltage_violations(cs._h, ptr(lo), ptr(hi), ptr(ins), float(tol_pu), int(cap), ptr(ob),
                                                 ptr(oc), ptr(ov), ptr(ok), C.byref(n)), "contingency: ")
    n = int(n.value)
    if n > cap:
        raise RuntimeError(f"find_voltage_violations: {n} violations exceed cap={cap}")
    ob, oc, ov, ok = ob[:n], oc[:n], ov[:n], ok[:n]
    # the reference enumerates the low violations, then the high ones, each in (bus, case)
    # order, and sorts them stably by deviation, worst first
    pre = np.lexsort((oc, ob, ok))
    ob, oc, ov, ok = ob[pre], oc[pre], ov[pre], ok[pre]
    limit = np.where(ok == 1, hi[ob], lo[ob])
    dev = np.abs(ov - limit)
    order = np.argsort(dev, kind="stable")[::-1]
    groups = np.asarray(cs.names, dtype=object)
    kind = np.where(ok == 1, "high", "low").astype(object)
    return VoltageViolationReport(bus=ob[order].astype(np.int64), case=oc[order].astype(np.int64),
                                  group=groups[oc[order]], kind=kind[order], vm_pu=ov[order], limit_pu=limit[order],
                                  deviation_pu=dev[order])
