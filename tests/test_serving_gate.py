"""The serving gate's order and width rules, on the host (no device).

The server asks the gate before each pass of a pool and each N-1 solve. The gate grants
shortest estimated work first, while the LU workgroups of the running grants fit its width;
a grant wider than the width runs when nothing else does; a worker that stops while it waits
holds nothing."""

import threading
import time

from p3smi.serving import _Gate, lu_width, work_estimate


def _waiter(gate, prio, w, log, stop=None):
    def run():
        ok = gate.acquire(prio, w, stop)
        log.append((prio, ok))

    th = threading.Thread(target=run, daemon=True)
    th.start()
    return th


def _settle():
    time.sleep(0.1)


def test_shortest_work_first_within_the_width():
    g = _Gate(10)
    assert g.acquire(5.0, 4)  # nothing runs: granted at once
    log = []
    big = _waiter(g, 3.0, 8, log)  # 4 + 8 > 10: waits
    _settle()
    assert log == []
    small = _waiter(g, 1.0, 2, log)  # shorter work, and 4 + 2 fits: passes the wider one
    small.join(2)
    assert log == [(1.0, True)]
    later = _waiter(g, 9.0, 1, log)  # fits, but longer work than the waiting head: waits behind it
    _settle()
    assert log == [(1.0, True)]
    g.release(4)  # 2 + 8 fits now; then 10 + 1 does not
    big.join(2)
    _settle()
    assert log == [(1.0, True), (3.0, True)]
    g.release(2)
    later.join(2)
    assert log == [(1.0, True), (3.0, True), (9.0, True)]
    g.release(8)
    g.release(1)
    assert g.used == 0 and g.running == 0 and g.grants == 4


def test_a_grant_wider_than_the_width_runs_alone():
    g = _Gate(10)
    assert g.acquire(1.0, 50)
    log = []
    th = _waiter(g, 0.5, 1, log)  # shorter, but 50 + 1 > 10
    _settle()
    assert log == []
    g.release(50)
    th.join(2)
    assert log == [(0.5, True)]
    g.release(1)
    assert g.used == 0 and g.running == 0


def test_a_stopped_waiter_holds_nothing_and_unblocks_the_rest():
    g = _Gate(10)
    assert g.acquire(1.0, 9)
    log, stop = [], threading.Event()
    head = _waiter(g, 2.0, 5, log, stop.is_set)  # waits: 9 + 5 > 10
    _settle()
    tail = _waiter(g, 3.0, 1, log)  # fits, but queued behind the head
    _settle()
    assert log == []
    stop.set()  # the head gives up: the next in line fits
    head.join(2)
    tail.join(2)
    assert sorted(log) == [(2.0, False), (3.0, True)]
    g.release(9)
    g.release(1)
    assert g.used == 0 and g.running == 0


def test_estimates():
    assert lu_width(1) == 1 and lu_width(256) == 256 and lu_width(257) == 5 and lu_width(4096) == 64
    assert lu_width(0) == 1
    # one case on a large grid outweighs a 64-case batch on a small one
    assert work_estimate(1_500_000, 1) > work_estimate(40_000, 64)
    assert work_estimate(1000, 10) < work_estimate(1000, 100)
