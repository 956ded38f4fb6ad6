"""CPU restatement of the reference's N-1 island and reference-bus rules (numpy/scipy).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's CPU leg, never by the product.

Restated from p3s/contingency/case_generator.py:
  * live branches (:316): a branch couples its two ends iff Y_ft != 0 or Y_tf != 0;
  * components after an outage (:355-377): scipy connected_components over the live,
    non-outaged branches (undirected);
  * reference resolution (:379-430): per component the lowest-index slack bus; else, with
    re-slacking, the lowest-|Z| generator bus, ties (np.isclose) to the lowest bus index;
    a bus is served iff its component has a reference;
  * the solver inputs (p3s/contingency/solver_cpp.py:119-150): re-slack references
    pinned at |V0| + 0j, unserved buses pinned at 1 + 0j, V set to NaN at unserved buses
    after the solve.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components


def components_after_outage(n, f, t, live, outaged):
    keep = np.asarray(live, dtype=bool).copy()
    if len(outaged):
        keep[np.asarray(outaged, dtype=np.int64)] = False
    fk, tk = np.asarray(f)[keep], np.asarray(t)[keep]
    rows = np.concatenate([fk, tk])
    cols = np.concatenate([tk, fk])
    adj = sp.coo_matrix((np.ones(rows.shape[0], dtype=np.int8), (rows, cols)), shape=(n, n)).tocsr()
    n_comp, labels = connected_components(adj, directed=False)
    return labels, n_comp


def resolve_references(labels, n_comp, ref, gen_bus, gen_z, reslack):
    """(served (n,) bool, ref_bus (n,) int64, pinned_refs list of bus indices)."""
    labels = np.asarray(labels)
    big = np.iinfo(np.int64).max
    comp_ref = np.full(n_comp, -1, dtype=np.int64)
    ref = np.asarray(ref, dtype=np.int64)
    if len(ref):
        tmp = np.full(n_comp, big, dtype=np.int64)
        np.minimum.at(tmp, labels[ref], ref)
        has = tmp != big
        comp_ref[has] = tmp[has]
    pinned = []
    gen_bus = np.asarray(gen_bus, dtype=np.int64)
    gen_z = np.asarray(gen_z, dtype=np.float64)
    if reslack and len(gen_bus):
        gcomp = labels[gen_bus]
        best_z = np.full(n_comp, np.inf)
        np.minimum.at(best_z, gcomp, gen_z)
        need = (comp_ref == -1) & np.isfinite(best_z)
        if need.any():
            is_best = np.isclose(gen_z, best_z[gcomp]) & need[gcomp]
            cand = np.where(is_best, gen_bus, big)
            chosen = np.full(n_comp, big, dtype=np.int64)
            np.minimum.at(chosen, gcomp, cand)
            picked = need & (chosen != big)
            comp_ref[picked] = chosen[picked]
            pinned = [int(chosen[c]) for c in np.nonzero(picked)[0]]
    ref_bus = comp_ref[labels]
    return ref_bus != -1, ref_bus, pinned


def cases(n, f, t, live, groups, ref, gen_bus, gen_z, reslack):
    """served (n, L) bool, ref_bus (n, L) int64, pinned refs per case."""
    L = len(groups)
    served = np.zeros((n, L), dtype=bool)
    ref_bus = np.zeros((n, L), dtype=np.int64)
    pinned = []
    for c, g in enumerate(groups):
        labels, nc = components_after_outage(n, f, t, live, g)
        s, r, p = resolve_references(labels, nc, ref, gen_bus, gen_z, reslack)
        served[:, c] = s
        ref_bus[:, c] = r
        pinned.append(p)
    return served, ref_bus, pinned


def solver_inputs(V_start, served, pinned):
    """Per-case V0 and pin (p3s/contingency/solver_cpp.py:119-134)."""
    n, L = served.shape
    V0 = np.repeat(np.asarray(V_start, dtype=np.complex128)[:, None], L, axis=1)
    pin = np.zeros((n, L), dtype=np.uint8)
    for c in range(L):
        for b in pinned[c]:
            V0[b, c] = np.abs(V0[b, c]) + 0.0j
            pin[b, c] = 1
        uns = ~served[:, c]
        pin[uns, c] = 1
        V0[uns, c] = 1.0 + 0.0j
    return V0, pin


def case_ybus(Yx_base, ypos, yval, group):
    """Base Ybus values minus the stamps of the outaged branches (case_generator.py:319-325)."""
    Yx = np.array(Yx_base, dtype=np.complex128)
    for k in group:
        for q in range(4):
            Yx[ypos[k, q]] -= yval[k, q]
    return Yx
