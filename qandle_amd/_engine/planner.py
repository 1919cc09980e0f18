"""Fusion planner: turns a gate list into passes of the fused kernel.

A pass owns K = R + TB local bit positions (always including positions
0..m-1 so tile loads are contiguous runs of 2^m amplitudes); inside a pass
the gates are packed into sub-passes that each hold R positions in registers.
What a gate needs:

* RX / RY on position p:   p in the sub-pass registers
* RZ on p:                  nothing (a per-thread phase when p is not a register bit)
* CNOT(c -> t):             t in the registers (c may be anywhere: its value is
                            then constant per thread)

Both levels are greedy over program order: positions are ranked by the index
of the first pending gate that needs them; a gate runs in the current
pass/sub-pass when its need is met and no earlier pending gate shares a
position with it; otherwise its positions are blocked for the rest of the scan
(later gates on them depend on it).  A pass closes when its estimated VALU
cost per amplitude reaches the budget, keeping it HBM-bound.

This replaces the reference's circuit splitter for execution purposes
(``src/qandle/splitter/grouping.py:30-64`` groups CNOTs into sub-circuits of
at most max_qubits qubits so that dense matrices stay small); here the unit is
a register/LDS-resident tile and nothing is ever densified.
"""
from __future__ import annotations

import dataclasses
import typing

import numpy as np

# op kinds (match qs_program.h for rotations)
RX, RY, RZ, CNOT = 0, 1, 2, 4
# gradient-only Pauli reductions (adjoint kernel): kind = PGRAD + Pauli (0 X, 1 Y, 2 Z)
PGRAD = 8
OP_RX, OP_RY, OP_RZ, OP_RZ_UNIFORM, OP_CNOT, OP_CNOT_UNIFORM, OP_PAULI_GRAD = 0, 1, 2, 3, 4, 5, 6
OP_PER_SAMPLE = 0x100
PAULI_UNIFORM = 0x10

HDR = 56
HDR_P, HDR_SWZ, HDR_GP, HDR_G, HDR_GG = 1, 8, 16, 24, 40


@dataclasses.dataclass(frozen=True)
class Gate:
    kind: int                 # RX, RY, RZ or CNOT
    pos: int                  # target bit position
    ctrl: int = -1            # control bit position (CNOT)
    slot: int = -1            # coefficient slot (rotations)

    @property
    def positions(self):
        return (self.pos,) if self.ctrl < 0 else (self.ctrl, self.pos)

    @property
    def need(self):
        """Positions that must be register bits for the gate to run."""
        if self.kind == RZ or self.kind == PGRAD + 2:
            return ()
        return (self.pos,)


@dataclasses.dataclass
class SubPass:
    P: typing.List[int]          # register positions (global), length R
    gates: typing.List[int]      # indices into the gate list


@dataclasses.dataclass
class Pass:
    L: typing.List[int]          # local positions (global, ascending), length K
    subs: typing.List[SubPass]
    gates: typing.List[int]


@dataclasses.dataclass
class Plan:
    n: int
    R: int
    TB: int
    passes: typing.List[Pass]

    @property
    def K(self):
        return self.R + self.TB


@dataclasses.dataclass
class CostModel:
    rot: float = 4.0
    rz_uniform: float = 4.0
    cnot: float = 1.0
    cnot_uniform: float = 2.0
    subpass: float = 3.0
    budget: float = 64.0        # VALU ops per amplitude per pass
    min_low: int = 5            # positions 0..min_low-1 are always local


def _gate_cost(g: Gate, in_regs: bool, cm: CostModel) -> float:
    if g.kind == CNOT:
        return cm.cnot
    if g.kind == RZ and not in_regs:
        return cm.rz_uniform
    return cm.rot


def _rank_positions(gates, pending, universe):
    """Positions of ``universe`` ordered by the first pending gate that needs them."""
    first = {}
    for idx in pending:
        for p in gates[idx].need:
            if p in universe and p not in first:
                first[p] = idx
    return sorted(first, key=lambda p: first[p])


def _scan(gates, pending, ok, budget=None, cost=None):
    """Greedy in-order selection; returns executed indices (in program order)."""
    blocked = set()
    done = []
    spent = 0.0
    for idx in pending:
        g = gates[idx]
        pos = g.positions
        if any(p in blocked for p in pos):
            blocked.update(pos)
            continue
        if ok(g):
            if budget is not None:
                c = cost(g)
                if done and spent + c > budget:
                    break
                spent += c
            done.append(idx)
        else:
            blocked.update(pos)
    return done


def _pack(gates, idxs, L, R, cm):
    """Sub-passes for the gates of one pass."""
    subs = []
    pending = list(idxs)
    Lset = set(L)
    while True:
        ranked = _rank_positions(gates, pending, Lset)
        P = ranked[:R]
        for p in L:  # fill with unused local positions
            if len(P) == R:
                break
            if p not in P:
                P.append(p)
        Pset = set(P)
        done = _scan(gates, pending, lambda g: all(p in Pset for p in g.need))
        subs.append(SubPass(sorted(P), done))
        ds = set(done)
        pending = [i for i in pending if i not in ds]
        if not pending:
            break
        if not done:
            raise RuntimeError("planner made no progress")
    return subs


def plan(n: int, gates: typing.List[Gate], R: int, TB: int, cm: CostModel = None) -> Plan:
    cm = cm or CostModel()
    K = R + TB
    if R > n:
        raise ValueError("register bits exceed qubit count")
    pending = list(range(len(gates)))
    passes = []
    if n < K:
        # the tile holds whole samples: one pass runs every gate
        L = list(range(K))
        subs = _pack(gates, pending, list(range(n)), R, cm) if pending else [SubPass(list(range(R)), [])]
        return Plan(n, R, TB, [Pass(L, subs, pending)])
    low = list(range(min(cm.min_low, K)))
    while True:
        ranked = [p for p in _rank_positions(gates, pending, set(range(n))) if p not in low]
        L = list(low) + ranked[: K - len(low)]
        for p in range(n):  # fill with the lowest unused positions
            if len(L) == K:
                break
            if p not in L:
                L.append(p)
        L = sorted(L)
        Lset = set(L)

        def ok(g):
            return all(p in Lset for p in g.need)

        def cost(g):
            return _gate_cost(g, g.kind != RZ or g.pos in Lset, cm)

        budget = cm.budget - cm.subpass * 2
        done = _scan(gates, pending, ok, budget, cost) if pending else []
        subs = _pack(gates, done, L, R, cm) if done else [SubPass(L[:R], [])]
        passes.append(Pass(L, subs, done))
        ds = set(done)
        pending = [i for i in pending if i not in ds]
        if not pending:
            break
        if not done:
            raise RuntimeError("planner made no progress")
    return Plan(n, R, TB, passes)


def swz(l: int) -> int:
    f = (l >> 5) ^ (l >> 9) ^ (l >> 13)
    return l ^ ((f & 15) << 1)


def encode_pass(p: Pass, gates, R: int, TB: int, coef_index, per_sample,
                reverse: bool = False, row_of=None) -> np.ndarray:
    """int32 words of one pass (layout in qs_program.h / qs_fused.hip).

    ``coef_index(slot)`` gives the float2 index of a rotation's coefficient
    (for a per-sample slot: the index of sample 0); ``per_sample(slot)`` says
    whether it advances with the sample index.  ``reverse`` encodes the pass
    for the adjoint kernel: sub-passes and the gates inside each in reverse
    order, and each rotation carries its gradient row ``row_of(gate)`` in
    word b.
    """
    K = R + TB
    L = p.L
    loc = {pos: j for j, pos in enumerate(L)}
    words = []
    for sp in (reversed(p.subs) if reverse else p.subs):
        Pl = [loc[q] for q in sp.P]                     # register local bits
        G = [j for j in range(K) if j not in Pl]        # thread-bit local bits
        hdr = [0] * HDR
        hdr[0] = len(sp.gates)
        for r, j in enumerate(Pl):
            hdr[HDR_P + r] = j
            hdr[HDR_SWZ + r] = swz(1 << j)
            hdr[HDR_GP + r] = L[j]
        for i, j in enumerate(G):
            hdr[HDR_G + i] = j
            hdr[HDR_GG + i] = L[j]
        words.extend(hdr)
        reg = {q: r for r, q in enumerate(sp.P)}
        for gi in (reversed(sp.gates) if reverse else sp.gates):
            g = gates[gi]
            if g.kind >= PGRAD:
                pauli = g.kind - PGRAD
                if g.pos in reg:
                    words.extend([OP_PAULI_GRAD, reg[g.pos], row_of(gi), pauli])
                else:
                    words.extend([OP_PAULI_GRAD, g.pos, row_of(gi), pauli | PAULI_UNIFORM])
                continue
            if g.kind == CNOT:
                if g.ctrl in reg:
                    words.extend([OP_CNOT, reg[g.pos], reg[g.ctrl], 0])
                else:
                    words.extend([OP_CNOT_UNIFORM, reg[g.pos], g.ctrl, 0])
                continue
            code = g.kind if g.pos in reg else OP_RZ_UNIFORM
            if code == OP_RZ_UNIFORM and g.kind != RZ:
                raise AssertionError("rotation scheduled off its register bit")
            if per_sample(g.slot):
                code |= OP_PER_SAMPLE
            a = reg[g.pos] if g.pos in reg else g.pos
            b = row_of(gi) if reverse else 0
            words.extend([code, a, b, int(coef_index(g.slot))])
    return np.asarray(words, dtype=np.int32)


def plan_stats(pl: Plan, gates, cm: CostModel = None):
    cm = cm or CostModel()
    out = []
    for p in pl.passes:
        Lset = set(p.L)
        cost = sum(_gate_cost(gates[i], gates[i].kind != RZ or gates[i].pos in Lset, cm) for i in p.gates)
        out.append({"gates": len(p.gates), "subpasses": len(p.subs), "cost": cost + cm.subpass * len(p.subs)})
    return out
