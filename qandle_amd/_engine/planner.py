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

# gate kinds (rotations match the QS_RX/RY/RZ codes of the C-ABI)
RX, RY, RZ, CNOT = 0, 1, 2, 4
# gradient-only Pauli reductions (adjoint kernel): kind = PGRAD + Pauli (0 X, 1 Y, 2 Z)
PGRAD = 8
CNOT_UNIFORM = 1      # CNOT word flag: control outside the registers (qs_program.h)

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
    lookahead: bool = False     # choose each pass's local positions by the score of the whole schedule


def _gate_cost(g: Gate, in_regs: bool, cm: CostModel) -> float:
    if g.kind == CNOT:
        return cm.cnot
    if g.kind == RZ and not in_regs:
        return cm.rz_uniform
    return cm.rot


def _rank_positions(gates, pending, universe, touch: bool = False):
    """Positions of ``universe`` ordered by the first pending gate that needs
    them (``touch``: that touches them, a CNOT's control included)."""
    first = {}
    for idx in pending:
        g = gates[idx]
        for p in (g.positions if touch else g.need):
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
        # register positions: the ones the next pending gates touch, so that a
        # CNOT whose control is also in registers is a free register swap
        ranked = _rank_positions(gates, pending, Lset, touch=True)
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
    """Passes (local position sets) and their sub-passes (register sets) for
    ``gates``.  Two greedy schedules are built, one choosing each pass's local
    positions by the positions the next pending gates touch, one taking per
    pass whichever of that and the positions they need (a CNOT's target only)
    runs more work; the plan with the lower score (4 per pass, a pass costing
    about four sub-passes of HBM time, plus 1 per sub-pass) is returned."""
    cm = cm or CostModel()
    K = R + TB
    if R > n:
        raise ValueError("register bits exceed qubit count")
    if n < K:
        # the tile holds whole samples: one pass runs every gate
        pending = list(range(len(gates)))
        L = list(range(K))
        subs = _pack(gates, pending, list(range(n)), R, cm) if pending else [SubPass(list(range(R)), [])]
        return Plan(n, R, TB, [Pass(L, subs, pending)])
    best = None
    for modes in ((True,), (False, True)):
        pl = _plan_passes(n, gates, R, TB, cm, modes)
        score = _score(pl)
        if best is None or score < best[0]:
            best = (score, pl)
    if cm.lookahead and gates:
        # also with one always-local position fewer (runs of 2^3 amplitudes, 64 bytes): the
        # windows then reach one position further, which can save passes (24-qubit ladder: 7
        # instead of 9 at K = 13, 3.03 vs 3.16-3.20 ms per forward; at 20 qubits 5 passes of 29
        # sub-passes score the same as 6 of 25 and measured the same, 2.36 vs 2.38 ms)
        for low in sorted({cm.min_low, max(cm.min_low - 1, 3)}, reverse=True):
            pl = _plan_lookahead(n, gates, R, TB, dataclasses.replace(cm, min_low=low))
            if _score(pl) < best[0]:
                best = (_score(pl), pl)
    return best[1]


def _score(pl: Plan) -> int:
    return 4 * len(pl.passes) + sum(len(p.subs) for p in pl.passes)


def _local_set(n: int, K: int, low, ranked):
    L = list(low) + [p for p in ranked if p not in low][: K - len(low)]
    for p in range(n):  # fill with the lowest unused positions
        if len(L) == K:
            break
        if p not in L:
            L.append(p)
    return sorted(L)


def _plan_lookahead(n: int, gates, R: int, TB: int, cm: CostModel) -> Plan:
    """Each pass takes, among candidate local sets (the touch and need
    rankings, and every window of K - min_low consecutive positions above the
    always-local ones), the one whose pass plus the greedy schedule of what it
    leaves scores lowest."""
    K = R + TB
    low = list(range(min(cm.min_low, K - 1)))
    budget = cm.budget - cm.subpass * 2
    high = [p for p in range(n) if p not in low]
    windows = [high[i:i + K - len(low)] for i in range(0, max(len(high) - (K - len(low)), 0) + 1)]
    pending = list(range(len(gates)))
    passes = []
    while pending:
        best = None
        cands = [_rank_positions(gates, pending, set(range(n)), t) for t in (True, False)] + windows
        seen = set()
        for ranked in cands:
            L = _local_set(n, K, low, ranked)
            if tuple(L) in seen:
                continue
            seen.add(tuple(L))
            Lset = set(L)

            def ok(g, Lset=Lset):
                return all(p in Lset for p in g.need)

            def cost(g, Lset=Lset):
                return _gate_cost(g, g.kind != RZ or g.pos in Lset, cm)

            done = _scan(gates, pending, ok, budget, cost)
            if not done:
                continue
            subs = _pack(gates, done, L, R, cm)
            ds = set(done)
            rest = [gates[i] for i in pending if i not in ds]
            score = 4 + len(subs) + (_score(_plan_passes(n, rest, R, TB, cm, (True,))) if rest else 0)
            if best is None or score < best[0]:
                best = (score, L, done, subs)
        if best is None:
            raise RuntimeError("planner made no progress")
        _, L, done, subs = best
        passes.append(Pass(L, subs, done))
        ds = set(done)
        pending = [i for i in pending if i not in ds]
    return Plan(n, R, TB, passes)


def _plan_passes(n: int, gates, R: int, TB: int, cm: CostModel, modes) -> Plan:
    """One greedy schedule; ``modes``: the local-set rankings tried per pass
    (True: by touch, False: by need), the one that runs more work is taken."""
    K = R + TB
    pending = list(range(len(gates)))
    passes = []
    low = list(range(min(cm.min_low, K - 1)))  # always leave room for a needed position
    budget = cm.budget - cm.subpass * 2
    while True:
        best = None
        for touch in modes:
            ranked = [p for p in _rank_positions(gates, pending, set(range(n)), touch) if p not in low]
            L = list(low) + ranked[: K - len(low)]
            for p in range(n):  # fill with the lowest unused positions
                if len(L) == K:
                    break
                if p not in L:
                    L.append(p)
            L = sorted(L)
            Lset = set(L)

            def ok(g, Lset=Lset):
                return all(p in Lset for p in g.need)

            def cost(g, Lset=Lset):
                return _gate_cost(g, g.kind != RZ or g.pos in Lset, cm)

            done = _scan(gates, pending, ok, budget, cost) if pending else []
            work = sum(cost(gates[i]) for i in done)
            if best is None or work > best[0]:
                best = (work, L, done)
        _, L, done = best
        subs = _pack(gates, done, L, R, cm) if done else [SubPass(L[:R], [])]
        passes.append(Pass(L, subs, done))
        ds = set(done)
        pending = [i for i in pending if i not in ds]
        if not pending:
            break
        if not done:
            raise RuntimeError("planner made no progress")
    return Plan(n, R, TB, passes)


def swz(l: int, mask: int = 31) -> int:
    """LDS slot of tile element l (same XOR-linear map as the kernels' swz;
    ``mask`` 30 keeps bit 0, so an amplitude pair stays in order in one 16-byte slot)."""
    f = (l >> 5) ^ (l >> 10) ^ (l >> 15)
    return l ^ (f & mask)


def plan_stats(pl: Plan, gates, cm: CostModel = None):
    cm = cm or CostModel()
    out = []
    for p in pl.passes:
        Lset = set(p.L)
        cost = sum(_gate_cost(gates[i], gates[i].kind != RZ or gates[i].pos in Lset, cm) for i in p.gates)
        out.append({"gates": len(p.gates), "subpasses": len(p.subs), "cost": cost + cm.subpass * len(p.subs)})
    return out


# ----------------------------------------------------------------------------
# word encoding (layout in qs_program.h)
# ----------------------------------------------------------------------------
#
# Single-qubit gates on different bits commute, and every RZ on a bit that is
# not a register bit commutes with everything else in the sub-pass (the only
# other use of such a bit is as a CNOT control).  So a sub-pass's gate list is
# rewritten as:
#   phases   : RZ gates on non-register bits -> one per-thread phase
#   segments : [per register bit, the product of its consecutive rotations]
#              followed by the CNOTs of that stretch, in order
# A rotation joins the current segment's unitary on its bit unless a CNOT of
# the segment already touched that bit (then a new segment starts).

SEG_NOUNAPPLY = 2
SEG_PER_SAMPLE = 1
@dataclasses.dataclass
class Segment:
    bits: dict                 # register index -> [gate indices, in application order]
    cnots: list                # gate indices


def segment_subpass(sp: SubPass, gates):
    reg = {q: r for r, q in enumerate(sp.P)}
    phases, segs = [], []
    cur = Segment({}, [])
    touched = set()
    for gi in sp.gates:
        g = gates[gi]
        if g.kind == CNOT:
            cur.cnots.append(gi)
            touched.add(reg[g.pos])
            if g.ctrl in reg:
                touched.add(reg[g.ctrl])
            continue
        if g.pos not in reg:  # RZ (or a Z gradient probe) on a non-register bit
            phases.append(gi)
            continue
        r = reg[g.pos]
        if r in touched:
            segs.append(cur)
            cur = Segment({}, [])
            touched = set()
        cur.bits.setdefault(r, []).append(gi)
    if cur.bits or cur.cnots or not segs:
        segs.append(cur)
    return phases, segs


class TableBuilder:
    """Allocates fused-unitary (U) and gradient-generator (Q) entries and records
    the device recipes that compose them (``qs_compose_tables``).

    Recipe words per U entry: u_index, span (B for a per-sample entry, else
    1), nconst, q_index, then per constituent (kind, cs_index, cs_per_sample);
    kind 0/1/2 = RX/RY/RZ, PGRAD + p = a gradient probe with Pauli p (identity
    rotation)."""

    def __init__(self, B: int):
        self.B = B
        self.nu = 0
        self.nq = 0
        self.words: list = []
        self.offsets: list = []

    def add(self, consts, B: int = None):
        ps = any(c[2] for c in consts)
        span = (self.B if B is None else B) if ps else 1
        u, q = self.nu, self.nq
        self.nu += span
        self.nq += span * len(consts)
        self.offsets.append(len(self.words))
        self.words.extend([u, span, len(consts), q])
        for c in consts:
            self.words.extend([int(c[0]), int(c[1]), 1 if c[2] else 0])
        return u, ps, q

    @property
    def count(self) -> int:
        return len(self.offsets)

    def recipes(self) -> np.ndarray:
        return np.asarray(self.words if self.words else [0], dtype=np.int32)


def thread_bits(K: int, TB: int, regs_list, keep_waves: bool):
    """Tile bits of the thread index of each sub-pass, lowest thread bit first.

    The bits that are not registers, ascending; with ``keep_waves`` the top
    thread bits (6 and up: the wave index of a 64-lane wave) are chosen so
    that consecutive sub-passes share them where their registers allow: each
    wave then holds the same part of the tile in both, and the kernels pass it
    on with a wave-level fence instead of a block barrier."""
    free = [[j for j in range(K) if j not in regs] for regs in regs_list]
    nw = TB - 6
    if not keep_waves or nw <= 0:
        return free
    out, prev = [], None
    for s, fr in enumerate(free):
        nxt = set(free[s + 1]) if s + 1 < len(free) else set(fr)
        if prev is not None and all(j in fr for j in prev):
            waves = prev
        else:
            both = [j for j in reversed(fr) if j in nxt]
            waves = (both + [j for j in reversed(fr) if j not in both])[:nw][::-1]
        out.append([j for j in fr if j not in waves] + list(waves))
        prev = waves
    return out


def encode_pass(p: Pass, gates, R: int, TB: int, coef_index, per_sample, tables: TableBuilder,
                reverse: bool = False, row_of=None, B: int = None, keep_waves: bool = False) -> np.ndarray:
    """int32 words of one pass in the layout of qs_program.h.

    ``coef_index(slot)`` is the float2 index of a rotation's (cos, sin) entry
    (for a per-sample slot: the entry of sample 0) and ``per_sample(slot)``
    says whether it advances with the sample.  Fused unitaries are allocated
    in ``tables``.  Forward segments are fixed records (head, one utab index
    per register bit) followed by their CNOT words.  ``reverse`` encodes for
    the adjoint kernel: sub-passes, segments and CNOTs in reverse order, each
    segment's CNOT words before its unitary entries (the order they are undone
    in), and every rotation's gradient row ``row_of(gate)``.  ``keep_waves``:
    thread_bits keeps the wave-index bits across sub-passes where it can (for
    tiles that hold part of one sample)."""
    K = R + TB
    L = p.L
    loc = {pos: j for j, pos in enumerate(L)}
    rows = (lambda gi: row_of(gi)) if (reverse and row_of is not None) else (lambda gi: -1)
    words = []
    subs = list(reversed(p.subs)) if reverse else p.subs
    Gs = thread_bits(K, TB, [[loc[q] for q in sp.P] for sp in subs], keep_waves)
    for sp, G in zip(subs, Gs):
        Pl = [loc[q] for q in sp.P]
        phases, segs = segment_subpass(sp, gates)
        if reverse:
            segs = list(reversed(segs))
        hdr = [0] * HDR
        hdr[0] = len(segs)
        hdr[6] = len(phases)
        for r, j in enumerate(Pl):
            hdr[HDR_P + r] = j
            hdr[HDR_SWZ + r] = swz(1 << j)
            hdr[HDR_GP + r] = L[j]
        for i, j in enumerate(G):
            hdr[HDR_G + i] = j
            hdr[HDR_GG + i] = L[j]
        words.extend(hdr)
        reg = {q: r for r, q in enumerate(sp.P)}
        for gi in phases:
            g = gates[gi]
            if g.kind >= PGRAD:  # gradient probe only (Z on a non-register bit)
                words.extend([g.pos, 0, SEG_NOUNAPPLY, rows(gi)])
            else:
                words.extend([g.pos, int(coef_index(g.slot)), SEG_PER_SAMPLE if per_sample(g.slot) else 0, rows(gi)])
        for seg in segs:
            cn = list(reversed(seg.cnots)) if reverse else seg.cnots
            umask = psmask = probemask = 0
            entries = {}
            for r in sorted(seg.bits):
                gl = seg.bits[r]
                consts = []
                probe = False
                for gi in gl:
                    g = gates[gi]
                    if g.kind >= PGRAD:
                        probe = True
                        consts.append((g.kind, 0, False))
                    else:
                        consts.append((g.kind, int(coef_index(g.slot)), bool(per_sample(g.slot))))
                u, ps, q = tables.add(consts, B)
                umask |= 1 << r
                psmask |= (1 << r) if ps else 0
                probemask |= (1 << r) if probe else 0
                entries[r] = (u, q, gl)
            words.append(umask | (len(cn) << 8) | (psmask << 16) | (probemask << 24))
            cwords = [_cnot_word(gates[gi], reg) for gi in cn]
            if reverse:
                words.extend(cwords)
                for r in sorted(entries):
                    u, q, gl = entries[r]
                    words.extend([u, len(gl), q])
                    words.extend(rows(gi) for gi in gl)
            else:
                words.extend(entries[r][0] if r in entries else 0 for r in range(R))
                words.extend(cwords)
    return np.asarray(words, dtype=np.int32)


def _cnot_word(g: Gate, reg) -> int:
    if g.ctrl in reg:
        return (reg[g.pos] << 8) | (reg[g.ctrl] << 16)
    return CNOT_UNIFORM | (reg[g.pos] << 8) | (g.ctrl << 16)


@dataclasses.dataclass
class DecodedSub:
    P: list          # register bit r -> tile-local bit
    G: list          # thread bit i -> tile-local bit
    GP: list         # register bit r -> global position
    GG: list         # thread bit i -> global position
    phases: list     # (global position, cs index, flags)
    segments: list   # ([(r, utab index, per_sample)], [(uniform, target register, control)])


def decode_pass(words, R: int, TB: int):
    """Forward words of one pass -> sub-pass structures (the layout of qs_program.h)."""
    w = [int(x) for x in words]
    pos, out = 0, []
    while pos < len(w):
        hdr = w[pos: pos + HDR]
        nseg, nph = hdr[0], hdr[6]
        sub = DecodedSub([hdr[HDR_P + r] for r in range(R)], [hdr[HDR_G + i] for i in range(TB)],
                         [hdr[HDR_GP + r] for r in range(R)], [hdr[HDR_GG + i] for i in range(TB)], [], [])
        p = pos + HDR
        for _ in range(nph):
            sub.phases.append((w[p], w[p + 1], w[p + 2]))
            p += 4
        for _ in range(nseg):
            head = w[p]
            umask, ncnot, psmask = head & 0xFF, (head >> 8) & 0xFF, (head >> 16) & 0xFF
            idx = w[p + 1: p + 1 + R]
            p += 1 + R
            bits = [(r, idx[r], bool((psmask >> r) & 1)) for r in range(R) if (umask >> r) & 1]
            cn = []
            for cw in w[p: p + ncnot]:
                cn.append((bool(cw & CNOT_UNIFORM), (cw >> 8) & 0xFF, cw >> 16))
            p += ncnot
            sub.segments.append((bits, cn))
        out.append(sub)
        pos = p
    if pos != len(w):
        raise ValueError("malformed pass words")
    return out


def decode_adjoint_pass(words, R: int, TB: int):
    """Adjoint (reverse-encoded) words of one pass -> sub-pass structures.

    Segments are (cnots, bits) with cnots = [(uniform, target, control)] and
    bits = [(r, utab index, per_sample, probe, qtab index, [rows])]; phases
    are (position, cs index, flags, row)."""
    w = [int(x) for x in words]
    pos, out = 0, []
    while pos < len(w):
        hdr = w[pos: pos + HDR]
        nseg, nph = hdr[0], hdr[6]
        sub = DecodedSub([hdr[HDR_P + r] for r in range(R)], [hdr[HDR_G + i] for i in range(TB)],
                         [hdr[HDR_GP + r] for r in range(R)], [hdr[HDR_GG + i] for i in range(TB)], [], [])
        p = pos + HDR
        for _ in range(nph):
            sub.phases.append((w[p], w[p + 1], w[p + 2], w[p + 3]))
            p += 4
        for _ in range(nseg):
            head = w[p]
            umask, ncnot = head & 0xFF, (head >> 8) & 0xFF
            psmask, probemask = (head >> 16) & 0xFF, (head >> 24) & 0x7F
            p += 1
            cn = [(bool(cw & CNOT_UNIFORM), (cw >> 8) & 0xFF, cw >> 16) for cw in w[p: p + ncnot]]
            p += ncnot
            bits = []
            for r in range(R):
                if (umask >> r) & 1:
                    u, nc, q = w[p], w[p + 1], w[p + 2]
                    rows = w[p + 3: p + 3 + nc]
                    bits.append((r, u, bool((psmask >> r) & 1), bool((probemask >> r) & 1), q, rows))
                    p += 3 + nc
            sub.segments.append((cn, bits))
        out.append(sub)
        pos = p
    if pos != len(w):
        raise ValueError("malformed adjoint pass words")
    return out
