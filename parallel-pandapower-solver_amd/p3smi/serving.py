"""Continuous-batching power-flow server on one GPU.

A request is (topology, Sbus (n, B), V0 (n,) or (n, B)): B cases on one grid. The server
keeps, per topology (keyed by a hash of Yp, Yj, Yx, pv, pq), a resident pool: the solver
plan, Ybus and a device batch of slots. Cases of every request queued for a topology are
loaded into free slots while the pool's other slots keep iterating; every pass advances
all loaded cases by one Newton iteration; finished slots are fetched, their results
scattered into their requests, and the slots refilled. A request completes when all its
cases have. Each pool is driven by its own host thread on its own HIP stream, so passes of
different topologies overlap on the device.

Admission is by device memory: a pool's slot count is its share of the memory budget
(``mem_fraction`` of the free HBM at start) divided by the footprint of one case; a new
topology that does not fit evicts idle pools first and otherwise waits.

No arithmetic mixes slots, so each case's result is bit-identical to solving it alone.
Reported per request: submit-to-completion latency (p50/p99 over requests in ``stats``).
"""

from __future__ import annotations

import ctypes as C
import hashlib
import threading
import time
from collections import OrderedDict, deque

import numpy as np

from . import _lib
from .nr_klu import Solver, _prec

_STATUS_ACTIVE, _STATUS_NEW = 0, 7


def topology_key(Yp, Yj, Yx, pv, pq) -> str:
    h = hashlib.sha1()
    for a, dt in ((Yp, np.int32), (Yj, np.int32), (Yx, np.complex128), (pv, np.int32), (pq, np.int32)):
        h.update(np.ascontiguousarray(a, dtype=dt).tobytes())
        h.update(b"|")
    return h.hexdigest()


def case_footprint(info: dict) -> int:
    """Device bytes one slot of a pool holds (state, Jacobian, factors, inputs, scratch)."""
    n, m, nnzj, nnzlu = info["n"], info["m"], info["nnz_j"], info["nnz_lu"]
    return int(8 * (6 * n + nnzj + nnzlu + 2 * m) + 16 * 2 * n + 8 * ((n + 63) // 64) + 128)


class Request:
    def __init__(self, key, S, V0):
        self.key = key
        self.S = np.ascontiguousarray(S, dtype=np.complex128)
        self.B = int(self.S.shape[1])
        n = self.S.shape[0]
        V0 = np.ascontiguousarray(V0, dtype=np.complex128)
        self.V0 = np.repeat(V0[:, None], self.B, axis=1) if V0.ndim == 1 else V0
        self.V = np.empty((n, self.B), dtype=np.complex128)
        self.iterations = np.empty(self.B, dtype=np.int32)
        self.converged = np.empty(self.B, dtype=bool)
        self.mismatch = np.empty(self.B, dtype=np.float64)
        self.remaining = self.B
        self.t_submit = time.perf_counter()
        self.t_done = None
        self._done = threading.Event()
        self.error = None

    @property
    def latency(self) -> float:
        return (self.t_done or time.perf_counter()) - self.t_submit

    def wait(self, timeout=None) -> bool:
        ok = self._done.wait(timeout)
        if self.error:
            raise self.error
        return ok

    def result(self) -> dict:
        self.wait()
        return {"V": self.V, "iterations": self.iterations, "converged": self.converged, "mismatch": self.mismatch}


class _Pool:
    def __init__(self, server, key, solver: Solver, slots: int):
        self.server = server
        self.key = key
        self.solver = solver
        self.n = solver.n
        self.capacity = slots
        self.L = _lib.lib()
        self.bh = C.c_void_p()
        _lib.check(self.L.p3smi_batch_create(solver._h, slots, C.byref(self.bh)))
        _lib.check(self.L.p3smi_batch_slots_init(self.bh))
        self.free = deque(range(slots))
        self.owner = {}  # slot -> (request, column)
        self.queue = deque()  # (request, next column)
        self.cv = threading.Condition()
        self.last_used = time.perf_counter()
        self.thread = threading.Thread(target=self._run, daemon=True)
        self.stop = False
        self.busy = False
        self.thread.start()

    def enqueue(self, req):
        with self.cv:
            self.queue.append([req, 0])
            self.cv.notify()

    def idle(self) -> bool:
        with self.cv:
            return not self.queue and not self.owner and not self.busy

    def _admit(self):
        """Move queued cases into free slots (one load call per request chunk)."""
        while True:
            with self.cv:
                if not self.queue or not self.free:
                    return
                item = self.queue[0]
                req, col = item
                k = min(len(self.free), req.B - col)
                slots = np.array([self.free.popleft() for _ in range(k)], dtype=np.int32)
                item[1] = col + k
                if item[1] >= req.B:
                    self.queue.popleft()
            S = np.ascontiguousarray(req.S[:, col:col + k])
            V0 = np.ascontiguousarray(req.V0[:, col:col + k])
            try:
                _lib.check(self.L.p3smi_batch_load(self.bh, k, _lib.ptr(slots), _lib.ptr(S), _lib.ptr(V0)))
            except Exception as e:  # bad inputs fail the request, the pool goes on
                req.error = e
                req.t_done = time.perf_counter()
                req._done.set()
                with self.cv:
                    self.free.extend(slots.tolist())
                continue
            with self.cv:
                for j, s in enumerate(slots.tolist()):
                    self.owner[s] = (req, col + j)

    def _harvest(self):
        buf = np.empty(self.capacity, dtype=np.int32)
        nfin = C.c_int32(0)
        _lib.check(self.L.p3smi_batch_finished(self.bh, _lib.ptr(buf), C.byref(nfin)))
        k = int(nfin.value)
        if k == 0:
            return
        slots = np.ascontiguousarray(buf[:k])
        V = np.empty((self.n, k), dtype=np.complex128)
        it = np.empty(k, dtype=np.int32)
        cv = np.empty(k, dtype=np.uint8)
        mis = np.empty(k, dtype=np.float64)
        _lib.check(self.L.p3smi_batch_fetch(self.bh, k, _lib.ptr(slots), _lib.ptr(V), _lib.ptr(it), _lib.ptr(cv),
                                            _lib.ptr(mis)))
        done = []
        with self.cv:
            for j, s in enumerate(slots.tolist()):
                req, col = self.owner.pop(s)
                req.V[:, col] = V[:, j]
                req.iterations[col] = it[j]
                req.converged[col] = bool(cv[j])
                req.mismatch[col] = mis[j]
                req.remaining -= 1
                if req.remaining == 0:
                    done.append(req)
                self.free.append(s)
        now = time.perf_counter()
        for req in done:
            req.t_done = now
            req._done.set()
            self.server._finished(req)

    def _run(self):
        srv = self.server
        counts = np.zeros(8, dtype=np.int32)
        while True:
            with self.cv:
                while not self.stop and not self.queue and not self.owner:
                    self.busy = False
                    self.cv.wait(0.05)
                if self.stop:
                    return
                self.busy = True
            self._admit()
            _lib.check(self.L.p3smi_batch_pass(self.bh, srv.max_iter, srv.tol, int(srv.line_search), srv.prec,
                                               _lib.ptr(counts)))
            self._harvest()
            self.last_used = time.perf_counter()

    def close(self):
        with self.cv:
            self.stop = True
            self.cv.notify()
        self.thread.join()
        self.L.p3smi_batch_free(self.bh)
        self.bh = None


class Server:
    """Continuous-batching power-flow server (see module docstring)."""

    def __init__(self, tol: float = 1e-8, max_iter: int = 20, precision: str = "double", line_search: bool = True,
                 mem_fraction: float = 0.6, max_slots: int = 4096, mem_budget: int | None = None):
        self.tol = float(tol)
        self.max_iter = int(max_iter)
        self.prec = _prec(precision)
        self.line_search = bool(line_search)
        self.max_slots = int(max_slots)
        if mem_budget is None:
            free = C.c_size_t(0)
            total = C.c_size_t(0)
            hip = C.CDLL("libamdhip64.so")
            if hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
                raise RuntimeError("serving: hipMemGetInfo failed (no device?)")
            mem_budget = int(free.value * mem_fraction)
        self.budget = int(mem_budget)
        self.used = 0
        self.pools: OrderedDict[str, _Pool] = OrderedDict()
        self.solvers: dict = {}
        self.lock = threading.Lock()
        self.finished: list = []
        self.pending: deque = deque()  # requests waiting for memory

    # ------------------------------------------------------------------ API
    def submit(self, Yp, Yj, Yx, pv, pq, S, V0) -> Request:
        key = topology_key(Yp, Yj, Yx, pv, pq)
        req = Request(key, S, V0)
        with self.lock:
            pool = self.pools.get(key)
            if pool is None:
                pool = self._make_pool(key, Yp, Yj, Yx, pv, pq)
            if pool is None:
                self.pending.append((req, (Yp, Yj, Yx, pv, pq)))
                return req
            self.pools.move_to_end(key)
        pool.enqueue(req)
        return req

    def _make_pool(self, key, Yp, Yj, Yx, pv, pq):
        solver = self.solvers.get(key)
        if solver is None:
            solver = Solver(Yp, Yj, Yx, pv, pq)
            from . import autotune

            autotune.apply_cached(solver, self.max_slots)
            self.solvers[key] = solver
        per = case_footprint(solver.info)
        want = self.max_slots * per
        while self.used + min(want, 64 * per) > self.budget:
            victim = next((k for k, p in self.pools.items() if p.idle()), None)
            if victim is None:
                return None
            self._evict(victim)
        slots = int(min(self.max_slots, (self.budget - self.used) // per))
        slots = max(64, slots // 64 * 64) if slots >= 64 else slots
        if slots <= 0:
            return None
        pool = _Pool(self, key, solver, slots)
        pool.bytes = slots * per
        self.used += pool.bytes
        self.pools[key] = pool
        return pool

    def _evict(self, key):
        pool = self.pools.pop(key)
        pool.close()
        self.used -= pool.bytes

    def _finished(self, req):
        with self.lock:
            self.finished.append(req)
            waiting = list(self.pending)
            self.pending.clear()
        for req2, topo in waiting:  # retry admission of requests that waited for memory
            with self.lock:
                pool = self.pools.get(req2.key) or self._make_pool(req2.key, *topo)
                if pool is None:
                    self.pending.append((req2, topo))
                    continue
            pool.enqueue(req2)

    def wait_all(self, requests, timeout=None):
        t_end = None if timeout is None else time.perf_counter() + timeout
        for r in requests:
            left = None if t_end is None else max(0.0, t_end - time.perf_counter())
            if not r.wait(left):
                raise TimeoutError("serving: request not finished in time")

    def stats(self) -> dict:
        with self.lock:
            lat = np.array([r.latency for r in self.finished]) if self.finished else np.zeros(1)
            cases = sum(r.B for r in self.finished)
            conv = sum(int(r.converged.sum()) for r in self.finished)
        return {"requests": len(self.finished), "cases": cases, "converged": conv,
                "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
                "pools": len(self.pools), "mem_used": self.used, "mem_budget": self.budget}

    def close(self):
        with self.lock:
            keys = list(self.pools)
        for k in keys:
            self._evict(k)
