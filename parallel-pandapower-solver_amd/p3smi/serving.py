"""Continuous-batching power-flow server on one GPU.

A request is (topology, Sbus (n, B), V0 (n,) or (n, B)): B cases on one grid. The server
keeps, per topology (keyed by a hash of Yp, Yj, Yx, pv, pq), a resident pool: the solver
plan, Ybus and a device batch of slots. Cases of every request queued for a topology are
loaded into free slots while the pool's other slots keep iterating; every pass moves each
loaded case one stage on (its start point, a Newton step, or a line-search trial);
finished slots are fetched, their results scattered into their requests, and the slots
refilled. A request completes when all its
cases have. Each pool is driven by its own host thread on its own HIP stream, so passes of
different topologies overlap on the device.

Admission is by device memory: a pool gets as many slots as its first request has cases
(a multiple of 64, at most ``max_slots``), within what is left of the memory budget
(``mem_fraction`` of the free HBM at start) at the footprint of one case. A topology that
does not fit evicts idle pools first and otherwise waits; an idle pool too small for a new
request is rebuilt larger. ``submit`` queues a request on its live pool itself; a request
that needs a pool built or grown, or memory freed, goes to a routing thread, so that the
caller never waits on a device allocation.

N-1 requests (``submit_contingency``: a subset of branch-outage cases of a topology) run on
that topology's resident outage set (islands, then Newton on the base Ybus minus each case's
stamps, on the device), on a
solver handle and stream of their own with one worker thread per topology, so they overlap
with the pools' passes and with the N-1 work of other topologies.

Scheduling (``priority=True``, the default): before each pass of a pool, and before each N-1
solve, the worker asks a gate for the device. The gate grants in order of the estimated
remaining work of the smallest request the worker holds (factor entries x (64 + cases still
iterating)), shortest first, while the LU workgroups of the running work fit ``gpu_width``
(one workgroup per case for few-case launches, per 64 cases otherwise; one grant always runs;
``max_grants`` caps the grants running at once). Under a load above what the GPU serves,
small requests then pass the large ones instead of sharing the device with all of them.
``priority=False`` lets every worker run at once. ``age_scale`` (seconds) divides a request's
key by 1 + age / age_scale, so that long requests are not held back indefinitely.

No arithmetic mixes slots, and every LU launch shape sums in one order, so each case's
result is bit-identical to solving it alone, whatever the order of the passes. Reported per
request: submit-to-completion latency (p50/p99 over requests in ``stats``).
"""

from __future__ import annotations

import ctypes as C
import hashlib
import heapq
import itertools
import threading
import time
from collections import OrderedDict, deque

import numpy as np

from . import _lib
from .nr_klu import Solver, _prec



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


def _round64(k: int) -> int:
    return (int(k) + 63) // 64 * 64


def work_estimate(nnz_lu: int, cases: int) -> float:
    """Relative device time of the rest of a request: a factorization's operand count times
    (64 + cases), since up to a few dozen cases a pass costs about what one case costs (the
    LU is latency bound there) and beyond that it grows with the cases."""
    return float(nnz_lu) * (64 + int(cases))


def lu_width(cases: int, few_max: int = 256) -> int:
    """LU workgroups of a pass over `cases` iterating cases: one per case in the few-case
    shape, one per 64 cases in the batch shape."""
    cases = int(cases)
    return max(1, cases if cases <= few_max else (cases + 63) // 64)


class _Gate:
    """Grants the device to workers shortest estimated work first, while the workgroups of the
    grants running fit `width` (a grant wider than `width` runs when nothing else does)."""

    def __init__(self, width: int, max_grants: int | None = None):
        self.width = int(width)
        self.max_grants = max_grants
        self.lock = threading.Lock()
        self.used = 0
        self.running = 0
        self.heap = []  # (priority, seq, ticket)
        self.seq = itertools.count()
        self.grants = 0

    def _grant(self):
        while self.heap:
            ticket = self.heap[0][2]
            if ticket["cancelled"]:
                heapq.heappop(self.heap)
                continue
            if self.running and (self.used + ticket["w"] > self.width
                                 or (self.max_grants and self.running >= self.max_grants)):
                return
            heapq.heappop(self.heap)
            self.used += ticket["w"]
            self.running += 1
            self.grants += 1
            ticket["ev"].set()

    def acquire(self, priority: float, w: int, stop=None) -> bool:
        """Wait for a grant; False (nothing held) if `stop()` turns true meanwhile."""
        ticket = {"w": int(w), "ev": threading.Event(), "cancelled": False}
        with self.lock:
            heapq.heappush(self.heap, (float(priority), next(self.seq), ticket))
            self._grant()
        while not ticket["ev"].wait(0.05):
            if stop is not None and stop():
                with self.lock:
                    if ticket["ev"].is_set():  # granted meanwhile: give it back
                        self.used -= ticket["w"]
                        self.running -= 1
                    ticket["cancelled"] = True
                    self._grant()
                return False
        return True

    def release(self, w: int):
        with self.lock:
            self.used -= int(w)
            self.running -= 1
            self._grant()


class _NoGate:
    def acquire(self, priority, w, stop=None) -> bool:
        return True

    def release(self, w):
        pass


class Request:
    def __init__(self, key, S, V0, cases=None):
        self.key = key
        if cases is None:  # time series: S (n, B), V0 (n,) or (n, B)
            self.S = np.ascontiguousarray(S, dtype=np.complex128)
            self.B = int(self.S.shape[1])
            self.V0 = np.ascontiguousarray(V0, dtype=np.complex128)  # (n,) is shared by every case
        else:  # N-1 subset: S (n,), V0 (n,) start, outage case indices
            self.S = np.ascontiguousarray(S, dtype=np.complex128).ravel()
            self.V0 = np.ascontiguousarray(V0, dtype=np.complex128).ravel()
            self.cases = np.ascontiguousarray(cases, dtype=np.int64).ravel()
            self.B = int(self.cases.size)
        n = self.S.shape[0]
        self.served = None
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
        out = {"V": self.V, "iterations": self.iterations, "converged": self.converged, "mismatch": self.mismatch}
        if self.served is not None:
            out["served"] = self.served
        return out


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
        _lib.check(self.L.p3smi_batch_set_graphs(self.bh, 1))  # FP64 passes replay captured graphs
        self.free = deque(range(slots))
        self.nnz_lu = int(solver.info["nnz_lu"])
        self.held = {}  # id -> request queued or in slots, not finished
        self.owner = {}  # slot -> (request, column)
        self.queue = deque()  # (request, next column)
        self.cv = threading.Condition()
        self.last_used = time.perf_counter()
        self.thread = threading.Thread(target=self._run, daemon=True)
        self.stop = False
        self.retired = False  # taken out of the server's table: accepts no more requests
        self.busy = False
        self.dead = None  # the exception that ended the worker, if one did
        self.thread.start()

    def enqueue(self, req) -> bool:
        """Queue a request; False if the pool's worker has died (the caller routes it anew)."""
        with self.cv:
            if self.dead is not None or self.retired:
                return False
            self.queue.append([req, 0])
            self.held[id(req)] = req
            self.cv.notify()
        return True

    def _claim(self):
        """(priority, LU workgroups) of the next pass, for the server's gate."""
        now = time.perf_counter()
        srv = self.server
        with self.cv:
            prio = min((srv.priority_of(self.nnz_lu, r.remaining, now - r.t_submit) for r in self.held.values()),
                       default=work_estimate(self.nnz_lu, 0))
            queued = sum(r.B - col for r, col in self.queue)
            return prio, lu_width(min(self.capacity, len(self.owner) + queued))

    def idle(self) -> bool:
        with self.cv:
            return self.dead is not None or (not self.queue and not self.owner and not self.busy)

    def retire(self) -> bool:
        """Stop accepting requests if the pool is idle (atomically with enqueue); False if busy."""
        with self.cv:
            if self.dead is None and (self.queue or self.owner or self.busy):
                return False
            self.retired = True
            return True

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
            if req.V0.ndim == 1:
                V0 = np.repeat(req.V0[:, None], k, axis=1)
            else:
                V0 = np.ascontiguousarray(req.V0[:, col:col + k])
            try:
                _lib.check(self.L.p3smi_batch_load(self.bh, k, _lib.ptr(slots), _lib.ptr(S), _lib.ptr(V0)))
            except Exception as e:  # bad inputs fail the request, the pool goes on
                req.error = e
                req.t_done = time.perf_counter()
                req._done.set()
                with self.cv:
                    self.free.extend(slots.tolist())
                    self.held.pop(id(req), None)
                continue
            with self.cv:
                for j, s in enumerate(slots.tolist()):
                    self.owner[s] = (req, col + j)

    def _harvest(self, counts) -> list:
        """Fetch the finished slots into their requests; returns the requests now complete.
        `counts` are the pass's slots per status: none converged, failed or stopped means
        nothing to fetch (and no device round trip)."""
        if counts[1] + counts[2] + counts[5] == 0:
            return []
        buf = np.empty(self.capacity, dtype=np.int32)
        nfin = C.c_int32(0)
        _lib.check(self.L.p3smi_batch_finished(self.bh, _lib.ptr(buf), C.byref(nfin)))
        k = int(nfin.value)
        if k == 0:
            return []
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
                    self.held.pop(id(req), None)
                self.free.append(s)
        now = time.perf_counter()
        for req in done:
            req.t_done = now
            req._done.set()
        return done

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
            prio, w = self._claim()
            if not srv.gate.acquire(prio, w, lambda: self.stop):
                return
            try:
                self._admit()
                _lib.check(self.L.p3smi_batch_pass(self.bh, srv.max_iter, srv.tol, int(srv.line_search), srv.prec,
                                                   _lib.ptr(counts)))
                done = self._harvest(counts)
            except Exception as e:  # the pool cannot go on: fail everything it holds
                srv.gate.release(w)
                self._fail_all(e)
                return
            srv.gate.release(w)
            for req in done:  # may admit requests that waited for memory: outside the grant
                srv._finished(req)
            self.last_used = time.perf_counter()

    def _fail_all(self, err):
        """End every request the pool holds (queued or in slots) with `err` and mark the pool
        dead: the server evicts it and routes later requests of its topology to a new pool."""
        with self.cv:
            self.dead = err
            self.busy = False
            reqs = {id(item[0]): item[0] for item in self.queue}
            reqs.update({id(r): r for r, _ in self.owner.values()})
            self.queue.clear()
            self.owner.clear()
            self.held.clear()
        now = time.perf_counter()
        for req in reqs.values():
            if req._done.is_set():
                continue
            req.error = err
            req.t_done = now
            req._done.set()
            self.server._finished(req)

    def close(self):
        with self.cv:
            self.stop = True
            self.cv.notify()
        self.thread.join()
        self.L.p3smi_batch_free(self.bh)
        self.bh = None
        # requests still held (a pool closed while busy) end with an error, not a hang
        if self.queue or self.owner:
            self._fail_all(RuntimeError("serving: pool closed"))


class Server:
    """Continuous-batching power-flow server (see module docstring)."""

    def __init__(self, tol: float = 1e-8, max_iter: int = 20, precision: str = "double", line_search: bool = True,
                 mem_fraction: float = 0.6, max_slots: int = 4096, mem_budget: int | None = None,
                 priority: bool = True, gpu_width: int = 512, age_scale: float | None = None,
                 max_grants: int | None = None):
        self.precision = precision
        self.tol = float(tol)
        self.max_iter = int(max_iter)
        self.prec = _prec(precision)
        self.line_search = bool(line_search)
        self.max_slots = int(max_slots)
        if mem_budget is None:
            free, _total = _lib.device_memory()
            mem_budget = int(free * mem_fraction)
        self.budget = int(mem_budget)
        self.used = 0
        self.pools: OrderedDict[str, _Pool] = OrderedDict()
        self.solvers: dict = {}
        self.lock = threading.Lock()
        self.finished: list = []
        self.pending: deque = deque()  # requests waiting for memory
        self.outage_sets: dict = {}  # topology key -> resident N-1 outage set
        self.n1_queues: dict = {}  # topology key -> deque of N-1 requests
        self.n1_threads: dict = {}
        self.n1_cv = threading.Condition()
        self.n1_stop = False
        self.gate = _Gate(gpu_width, max_grants) if priority else _NoGate()
        self.age_scale = None if age_scale is None else float(age_scale)
        # requests that need a pool built, grown or memory freed are routed by a thread of
        # their own, so that submit never waits on a device allocation
        self.route_q: deque = deque()
        self.route_cv = threading.Condition()
        self.route_stop = False
        self.router = threading.Thread(target=self._router, daemon=True)
        self.router.start()

    # ------------------------------------------------------------------ API
    def submit(self, Yp, Yj, Yx, pv, pq, S, V0) -> Request:
        key = topology_key(Yp, Yj, Yx, pv, pq)
        req = Request(key, S, V0)
        if not self._enqueue_resident(req):
            self._to_router(req, (Yp, Yj, Yx, pv, pq))
        return req

    def _enqueue_resident(self, req) -> bool:
        """Queue a request on its topology's live pool when no pool has to be built or grown."""
        want = min(self.max_slots, _round64(req.B))
        with self.lock:
            pool = self.pools.get(req.key)
            if pool is None or pool.dead is not None or (pool.capacity < want and pool.idle()):
                return False
            self.pools.move_to_end(req.key)
        return pool.enqueue(req)

    def _to_router(self, req, topo):
        with self.route_cv:
            self.route_q.append((req, topo))
            self.route_cv.notify()

    def _router(self):
        while True:
            with self.route_cv:
                while not self.route_stop and not self.route_q:
                    self.route_cv.wait(0.05)
                if self.route_stop:
                    return
                req, topo = self.route_q.popleft()
            try:
                self._route(req, topo)
            except Exception as e:  # e.g. a topology the solver rejects: that request fails alone
                req.error = e
                req.t_done = time.perf_counter()
                req._done.set()
                self._finished(req)

    def _route(self, req, topo):
        """Queue a request on its topology's pool, building or growing the pool first (routing
        thread; device allocations happen outside the server lock). A request that does not
        fit in memory waits for a finishing one."""
        want = min(self.max_slots, _round64(req.B))
        while True:
            victims = []
            with self.lock:
                pool = self.pools.get(req.key)
                if pool is not None and (pool.dead is not None or (pool.capacity < want and pool.retire())):
                    victims.append(self._unlist(req.key))  # a dead pool, or an idle one too small
                    pool = None
                if pool is not None:
                    self.pools.move_to_end(req.key)
            for v in victims:
                v.close()
            if pool is None:
                pool = self._make_pool(req.key, topo, want)
                if pool is None:
                    with self.lock:
                        self.pending.append((req, topo))
                    return
            if pool.enqueue(req):
                return

    def _unlist(self, key):
        """Take a pool out of the table and the memory count (lock held); the caller closes it."""
        pool = self.pools.pop(key)
        self.used -= pool.bytes
        return pool

    def _solver(self, key, Yp, Yj, Yx, pv, pq):
        solver = self.solvers.get(key)
        if solver is None:
            solver = Solver(Yp, Yj, Yx, pv, pq)
            self.solvers[key] = solver
        return solver

    def _make_pool(self, key, topo, want: int):
        """Build a pool of up to `want` slots, evicting idle pools for memory; None if even the
        smallest pool does not fit. Its bytes are reserved while it is built."""
        from . import autotune

        solver = self._solver(key, *topo)
        per = case_footprint(solver.info)
        victims = []
        with self.lock:
            while self.used + min(want, 64) * per > self.budget:
                victim = next((k for k, p in self.pools.items() if p.retire()), None)
                if victim is None:
                    break
                victims.append(self._unlist(victim))
            slots = int(min(want, (self.budget - self.used) // per))
            slots = slots // 64 * 64 if slots >= 64 else slots
            if slots > 0:
                self.used += slots * per
        for v in victims:
            v.close()
        if slots <= 0:
            return None
        try:
            autotune.apply_cached(solver, slots)
            pool = _Pool(self, key, solver, slots)
        except Exception:
            with self.lock:
                self.used -= slots * per
            raise
        pool.bytes = slots * per
        with self.lock:
            self.pools[key] = pool
        return pool

    # ------------------------------------------------------------------ N-1 requests
    def register_outages(self, Yp, Yj, Yx, pv, pq, f, t, stamps, groups, ref, gen_z=None, reslack=False) -> str:
        """Make the outage set of a topology resident: branches f -> t with stamps (yff, yft,
        ytf, ytt), outage groups (branch index lists), slack buses, generator |Z| per bus.
        Returns the topology key that submit_contingency takes."""
        from .contingency import ContingencySet

        key = topology_key(Yp, Yj, Yx, pv, pq)
        with self.lock:
            if key not in self.outage_sets:
                own = Solver(Yp, Yj, Yx, pv, pq)  # own handle and stream: runs beside the pool
                self.outage_sets[key] = ContingencySet(own, f, t, *stamps, groups, ref, gen_z, reslack)
                self.n1_queues[key] = deque()
                th = threading.Thread(target=self._n1_run, args=(key,), daemon=True)
                self.n1_threads[key] = th
                th.start()
        return key

    def submit_contingency(self, key: str, cases, S, V_start) -> Request:
        """Solve the outage cases `cases` (indices into the registered groups) of a topology."""
        if key not in self.outage_sets:
            raise KeyError("submit_contingency: register_outages first")
        req = Request(key, S, V_start, cases=cases)
        with self.n1_cv:
            self.n1_queues[key].append(req)
            self.n1_cv.notify_all()
        return req

    def _n1_run(self, key):
        from .contingency import ContingencySet

        queue = self.n1_queues[key]
        while True:
            with self.n1_cv:
                while not self.n1_stop and not queue:
                    self.n1_cv.wait(0.05)
                if self.n1_stop:
                    return
                req = queue.popleft()
            full = self.outage_sets[req.key]
            w = lu_width(req.B)
            try:
                sub = ContingencySet(full.solver, full.f, full.t, *full.yval.T, [full.groups[c] for c in req.cases],
                                     full.ref, full.gen_z, full.reslack)
                prio = self.priority_of(full.solver.info["nnz_lu"], req.B, time.perf_counter() - req.t_submit)
                if not self.gate.acquire(prio, w, lambda: self.n1_stop):
                    raise RuntimeError("serving: server closed")
                try:
                    r = sub.solve(req.S, req.V0, max_iter=self.max_iter, tol=self.tol, line_search=self.line_search,
                                  precision=self.precision)
                finally:
                    self.gate.release(w)
                req.V[:] = r.V
                req.iterations[:] = r.iterations
                req.converged[:] = r.converged
                req.mismatch[:] = r.mismatch
                req.served = r.served
            except Exception as e:  # a bad request fails alone
                req.error = e
            # the solve's workspace is not in the memory budget: give it back at once
            _lib.lib().p3smi_topology_release_cache(full.solver._h)
            req.remaining = 0
            req.t_done = time.perf_counter()
            req._done.set()
            self._finished(req)

    def priority_of(self, nnz_lu: int, cases: int, age: float) -> float:
        """The gate's key of a request (lower runs first): its estimated remaining work, divided
        by 1 + age / age_scale when an age scale is set, so that long requests waiting behind a
        stream of short ones are eventually served."""
        w = work_estimate(nnz_lu, cases)
        return w if self.age_scale is None else w / (1.0 + max(0.0, age) / self.age_scale)


    def _finished(self, req):
        with self.lock:
            self.finished.append(req)
            waiting = list(self.pending)
            self.pending.clear()
        for req2, topo in waiting:  # retry admission of requests that waited for memory
            self._to_router(req2, topo)

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
                "pools": len(self.pools), "mem_used": self.used, "mem_budget": self.budget,
                "gate_grants": getattr(self.gate, "grants", None)}

    def close(self):
        with self.route_cv:
            self.route_stop = True
            self.route_cv.notify_all()
        self.router.join()
        with self.n1_cv:
            self.n1_stop = True
            self.n1_cv.notify_all()
        for th in self.n1_threads.values():
            th.join()
        self.n1_threads.clear()
        with self.lock:
            pools = [self._unlist(k) for k in list(self.pools)]
        for pool in pools:
            pool.close()
        self.outage_sets.clear()
