"""Network ingestion from MATPOWER case files (version 2) to the solver's inputs.

The reference builds its Ybus from a pandapower net (p3s/NewtonPowerflowCpp.py:64-229,
p3s/models/*); real grids such as the pegase cases ship as MATPOWER files, so this module
reads those directly. A case file defines ``mpc.baseMVA`` and the matrices ``mpc.bus``,
``mpc.gen`` and ``mpc.branch`` (columns as in the MATPOWER manual, caseformat):

  bus:    BUS_I TYPE PD QD GS BS AREA VM VA BASE_KV ZONE VMAX VMIN
  gen:    GEN_BUS PG QG QMAX QMIN VG MBASE GEN_STATUS PMAX PMIN ...
  branch: F_BUS T_BUS BR_R BR_X BR_B RATE_A RATE_B RATE_C TAP SHIFT BR_STATUS ANGMIN ANGMAX

Ybus follows the standard pi model with the off-nominal tap on the from side: with
ys = 1 / (r + jx), tau = TAP (0 means 1) and theta = SHIFT (degrees),
  Ytt = ys + j b/2,  Yff = Ytt / tau^2,  Yft = -ys / (tau e^{-j theta}),
  Ytf = -ys / (tau e^{+j theta}),
plus the bus shunts (GS + j BS) / baseMVA on the diagonal; out-of-service branches are
left out. Bus types: 3 reference, 2 PV (if an in-service generator sits there, else PQ),
1 PQ, 4 isolated (left out of pv and pq, pinned to its start voltage). Sbus = (sum of the
in-service generators' PG + j QG - PD - j QD) / baseMVA; the start voltage is VM at VA,
with the generators' VG at reference and PV buses. Buses are numbered in file order.
"""

from __future__ import annotations

import re
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp


@dataclass
class Case:
    base_mva: float
    bus: np.ndarray
    gen: np.ndarray
    branch: np.ndarray
    # derived
    n: int
    bus_index: dict  # external bus number -> row
    f: np.ndarray
    t: np.ndarray
    stamps: tuple  # (Yff, Yft, Ytf, Ytt) of the in-service branches, in branch order
    shunt: np.ndarray  # (n,) complex bus shunt admittance, p.u.
    ref: np.ndarray
    pv: np.ndarray
    pq: np.ndarray
    isolated: np.ndarray
    sbus: np.ndarray
    v0: np.ndarray
    rate_a: np.ndarray  # MVA ratings of the in-service branches (0 = unlimited)

    def ybus(self) -> sp.csr_matrix:
        f, t = self.f, self.t
        yff, yft, ytf, ytt = self.stamps
        i = np.arange(self.n)
        rows = np.concatenate([f, f, t, t, i])
        cols = np.concatenate([f, t, f, t, i])
        vals = np.concatenate([yff, yft, ytf, ytt, self.shunt])
        Y = sp.coo_matrix((vals, (rows, cols)), shape=(self.n, self.n)).tocsr()
        Y.sum_duplicates()
        Y.sort_indices()
        return Y

    def csr(self):
        Y = self.ybus()
        return (np.ascontiguousarray(Y.indptr, dtype=np.int32), np.ascontiguousarray(Y.indices, dtype=np.int32),
                np.ascontiguousarray(Y.data, dtype=np.complex128))


_BLOCK = re.compile(r"mpc\.(\w+)\s*=\s*(\[.*?\]|[^;\n]+)\s*;", re.S)


def _matrix(text: str) -> np.ndarray:
    body = text.strip()[1:-1]
    rows = []
    for line in re.split(r"[;\n]", body):
        line = line.split("%", 1)[0].strip()
        if line:
            rows.append([float(x) for x in line.replace(",", " ").split()])
    width = max((len(r) for r in rows), default=0)
    if any(len(r) != width for r in rows):
        raise ValueError("MATPOWER case: ragged matrix rows")
    return np.array(rows, dtype=np.float64).reshape(len(rows), width)


def parse(text: str) -> Case:
    """Case from the text of a MATPOWER case file."""
    # drop comment lines first so that brackets inside comments do not confuse the blocks
    clean = "\n".join(line.split("%", 1)[0] for line in text.splitlines())
    blocks = {m.group(1): m.group(2) for m in _BLOCK.finditer(clean)}
    for key in ("baseMVA", "bus", "gen", "branch"):
        if key not in blocks:
            raise ValueError(f"MATPOWER case: mpc.{key} is missing")
    base = float(blocks["baseMVA"])
    bus, gen, branch = (_matrix(blocks[k]) for k in ("bus", "gen", "branch"))
    if bus.shape[1] < 13 or gen.shape[1] < 8 or branch.shape[1] < 11:
        raise ValueError("MATPOWER case: too few columns in bus / gen / branch")
    return build(base, bus, gen, branch)


def load(path: str) -> Case:
    with open(path) as fh:
        return parse(fh.read())


def build(base_mva: float, bus: np.ndarray, gen: np.ndarray, branch: np.ndarray) -> Case:
    n = bus.shape[0]
    ids = bus[:, 0].astype(np.int64)
    if len(set(ids.tolist())) != n:
        raise ValueError("MATPOWER case: duplicate bus numbers")
    index = {int(b): i for i, b in enumerate(ids)}

    def rows_of(col):
        try:
            return np.array([index[int(b)] for b in col], dtype=np.int64)
        except KeyError as e:
            raise ValueError(f"MATPOWER case: unknown bus {e.args[0]}") from None

    on = branch[:, 10] > 0
    br = branch[on]
    f, t = rows_of(br[:, 0]), rows_of(br[:, 1])
    ys = 1.0 / (br[:, 2] + 1j * br[:, 3])
    tau = np.where(br[:, 8] == 0.0, 1.0, br[:, 8])
    tap = tau * np.exp(1j * np.deg2rad(br[:, 9]))
    ytt = ys + 0.5j * br[:, 4]
    yff = ytt / (tap * np.conj(tap))
    yft = -ys / np.conj(tap)
    ytf = -ys / tap
    shunt = (bus[:, 4] + 1j * bus[:, 5]) / base_mva

    g_on = gen[:, 7] > 0
    g = gen[g_on]
    gb = rows_of(g[:, 0])
    btype = bus[:, 1].astype(np.int64)
    has_gen = np.zeros(n, bool)
    has_gen[gb] = True
    ref = np.nonzero(btype == 3)[0]
    if ref.size == 0:
        raise ValueError("MATPOWER case: no reference bus (type 3)")
    pv = np.nonzero((btype == 2) & has_gen)[0]
    pq = np.nonzero((btype == 1) | ((btype == 2) & ~has_gen))[0]
    isolated = np.nonzero(btype == 4)[0]

    sbus = -(bus[:, 2] + 1j * bus[:, 3]) / base_mva
    np.add.at(sbus, gb, (g[:, 1] + 1j * g[:, 2]) / base_mva)
    vm = bus[:, 7].astype(np.float64).copy()
    va = np.deg2rad(bus[:, 8])
    regulated = np.zeros(n, bool)
    regulated[ref] = True
    regulated[pv] = True
    for k in range(g.shape[0]):
        if regulated[gb[k]]:
            vm[gb[k]] = g[k, 5]
    v0 = vm * np.exp(1j * va)
    return Case(base_mva=base_mva, bus=bus, gen=gen, branch=branch, n=n, bus_index=index, f=f, t=t,
                stamps=(yff, yft, ytf, ytt), shunt=shunt, ref=ref, pv=pv, pq=pq, isolated=isolated,
                sbus=np.ascontiguousarray(sbus), v0=np.ascontiguousarray(v0), rate_a=br[:, 5].copy())
