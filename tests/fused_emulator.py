"""numpy emulation of the fused pass / adjoint kernels' v2 word semantics (test infrastructure).

Decodes exactly the int32 words and table recipes the planner sends to the GPU
(``qs_compose_tables``, ``qs_fused_pass``, ``qs_adjoint_pass``) and executes
them the way the kernels do: tile gather through the swizzled LDS index,
per-thread register groups, fused per-bit unitaries, per-thread phases,
uniform-condition CNOTs, product-state generation, measurement partials, and
for the adjoint the cross-density gradients.  Vectorised over the threads of a
tile.  It lets the planner and the encoder be checked against the CPU oracle
without a GPU.
"""
import numpy as np

from qandle_amd._engine import planner as PL

PAULI = [np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]], complex),
         np.array([[1, 0], [0, -1]], complex)]


def _dep(vals, positions):
    out = np.zeros_like(vals, dtype=np.int64)
    for i, p in enumerate(positions):
        out |= ((vals >> i) & 1) << p
    return out


def _rot(kind, c, s):
    if kind == 0:
        return np.array([[c, -1j * s], [-1j * s, c]])
    if kind == 1:
        return np.array([[c, -s], [s, c]], complex)
    if kind == 2:
        return np.array([[c - 1j * s, 0], [0, c + 1j * s]])
    return np.eye(2, dtype=complex)


def compose(recipes_words, B, cs, nu, nq):
    """numpy restatement of k_compose: returns utab (nu, 2, 2), qtab (nq, 2, 2)."""
    utab = np.zeros((max(nu, 1), 2, 2), complex)
    qtab = np.zeros((max(nq, 1), 2, 2), complex)
    w = list(recipes_words)
    i = 0
    while i < len(w) and nu:
        u, span, m, q = w[i: i + 4]
        consts = [w[i + 4 + 3 * k: i + 7 + 3 * k] for k in range(m)]
        i += 4 + 3 * m
        for b in range(span):
            Rs, Ps = [], []
            for kind, ci, cps in consts:
                if kind >= 8:
                    Rs.append(np.eye(2, dtype=complex))
                    Ps.append(kind - 8)
                else:
                    c, s = cs[ci + (b if cps else 0)]
                    Rs.append(_rot(kind, c, s))
                    Ps.append(kind)
            U = np.eye(2, dtype=complex)
            for R_ in Rs:
                U = R_ @ U
            utab[u + b] = U
            V = np.eye(2, dtype=complex)
            for k in range(m - 1, -1, -1):
                qtab[q + k * span + b] = V @ PAULI[Ps[k]] @ V.conj().T
                V = V @ Rs[k]
    return utab, qtab


def _cnot_fields(cw):
    """(uniform control?, target register, control register or global position)."""
    return bool(cw & PL.CNOT_UNIFORM), (cw >> 8) & 0xFF, cw >> 16


def _subpasses(words, R, adjoint=False):
    """Walks the words with one cursor in the order the kernel reads them.

    Returns per sub-pass (header, phases, segments); a segment is (bits, cnots)
    with bits = [(r, utab index, per-sample?, probe?, nconst, qtab index, rows)]
    for the masked register bits and cnots = [(uniform?, target, control)].
    The forward kernel reads a fixed record (head + R utab indices) then the
    CNOT words; the adjoint kernel the head, the CNOT words, then one variable
    entry per masked bit."""
    pos, out = 0, []
    while pos < len(words):
        hdr = words[pos: pos + PL.HDR]
        nseg, nph = int(hdr[0]), int(hdr[6])
        assert nseg >= 0 and nph >= 0
        p = pos + PL.HDR + 4 * nph
        segs = []
        for _ in range(nseg):
            head = int(words[p])
            p += 1
            umask, ncnot = head & 0xFF, (head >> 8) & 0xFF
            psmask, probemask = (head >> 16) & 0xFF, (head >> 24) & 0x7F
            assert umask < (1 << R) and (psmask | probemask) & ~umask == 0
            bits = []
            if adjoint:
                cn = [_cnot_fields(int(x)) for x in words[p: p + ncnot]]
                p += ncnot
                for r in range(R):
                    if (umask >> r) & 1:
                        ui, nc, qi = (int(x) for x in words[p: p + 3])
                        rows = [int(x) for x in words[p + 3: p + 3 + nc]]
                        assert len(rows) == nc
                        bits.append((r, ui, bool((psmask >> r) & 1), bool((probemask >> r) & 1), nc, qi, rows))
                        p += 3 + nc
            else:
                idx = [int(x) for x in words[p: p + R]]
                assert len(idx) == R
                p += R
                for r in range(R):
                    if (umask >> r) & 1:
                        bits.append((r, idx[r], bool((psmask >> r) & 1), False, 0, 0, []))
                    else:
                        assert idx[r] == 0
                cn = [_cnot_fields(int(x)) for x in words[p: p + ncnot]]
                p += ncnot
            assert len(cn) == ncnot
            segs.append((bits, cn))
        phases = [tuple(int(x) for x in words[pos + PL.HDR + 4 * k: pos + PL.HDR + 4 * k + 4]) for k in range(nph)]
        out.append((hdr, phases, segs))
        pos = p
    assert pos == len(words), "the walk must end exactly at the end of the words"
    return out


class _Geom:
    def __init__(self, B, n, R, TB, lpos):
        self.K = R + TB
        self.T, self.NR = 1 << TB, 1 << R
        self.multi = n < self.K
        self.total = B << n
        if self.multi:
            self.lpos = list(range(self.K))
            self.ntiles = -(-B // (1 << (self.K - n)))
            self.npos = []
        else:
            self.lpos = list(lpos)
            self.ntiles = B << (n - self.K)
            self.npos = [p for p in range(n) if p not in lpos]

    def base(self, t, n):
        if self.multi:
            return t << self.K
        nn = len(self.npos)
        tl = t & ((1 << nn) - 1)
        b = (t >> nn) << n
        for i in range(nn):
            if (tl >> i) & 1:
                b |= 1 << self.npos[i]
        return b


def _regs(hdr, R, TB, base, n, multi, T):
    tids = np.arange(T, dtype=np.int64)
    P = [int(hdr[PL.HDR_P + r]) for r in range(R)]
    SW = [int(hdr[PL.HDR_SWZ + r]) for r in range(R)]
    GP = [int(hdr[PL.HDR_GP + r]) for r in range(R)]
    G = [int(hdr[PL.HDR_G + i]) for i in range(TB)]
    GG = [int(hdr[PL.HDR_GG + i]) for i in range(TB)]
    gl = _dep(tids, G)
    gbase = base | _dep(tids, GG)
    sample = (gbase >> n) if multi else np.full(T, base >> n)
    swz = np.vectorize(PL.swz)
    addr = np.zeros((T, 1 << R), dtype=np.int64)
    addr[:, 0] = swz(gl)
    for j in range(1, 1 << R):
        addr[:, j] = addr[:, j & (j - 1)] ^ SW[(j & -j).bit_length() - 1]
    return P, GP, G, gbase, sample, addr, tids


def _apply_pair(v, r, M):
    for j in range(v.shape[1]):
        if (j >> r) & 1:
            continue
        j1 = j | (1 << r)
        a0, a1 = v[:, j].copy(), v[:, j1].copy()
        v[:, j] = M[:, 0, 0] * a0 + M[:, 0, 1] * a1
        v[:, j1] = M[:, 1, 0] * a0 + M[:, 1, 1] * a1


def _cnot(v, a, b):
    for j in range(v.shape[1]):
        if (j >> b) & 1 and not (j >> a) & 1:
            v[:, [j, j | (1 << a)]] = v[:, [j | (1 << a), j]]


def _swap_where(v, a, cond):
    for j in range(v.shape[1]):
        if not (j >> a) & 1:
            tmp = v[cond, j].copy()
            v[cond, j] = v[cond, j | (1 << a)]
            v[cond, j | (1 << a)] = tmp


def _phase(phases, cs, csample, gbase, inverse):
    z = np.ones(len(gbase), complex)
    for pos, ci, fl, _row in phases:
        if fl & PL.SEG_NOUNAPPLY:
            continue
        idx = ci + (csample if fl & PL.SEG_PER_SAMPLE else 0)
        c, s = cs[idx, 0], cs[idx, 1]
        sn = np.where((gbase >> pos) & 1, -s, s)
        if inverse:
            sn = -sn
        z = z * (c - 1j * sn)
    return z


def emulate_pass(state, src, B, n, R, TB, lpos, words, cs, utab, efac, measure):
    """Forward pass.  state/src: (B*2^n,) complex128 (src None -> generate from efac (B, n, 2))."""
    g = _Geom(B, n, R, TB, lpos)
    K, T, NR = g.K, g.T, g.NR
    spt_log = (K - n) if g.multi else 0
    mparts = np.zeros((g.ntiles << spt_log, K + 1))
    swz = np.vectorize(PL.swz)
    subs = _subpasses(words, R)
    for t in range(g.ntiles):
        base = g.base(t, n)
        tile = np.zeros(1 << K, dtype=np.complex128)
        ls = np.arange(1 << K, dtype=np.int64)
        gi_all = base | _dep(ls, g.lpos)
        ok = gi_all < g.total
        if src is not None:
            tile[swz(ls)] = np.where(ok, src[np.minimum(gi_all, g.total - 1)], 0)
        for si, (hdr, phases, segs) in enumerate(subs):
            P, GP, G, gbase, sample, addr, tids = _regs(hdr, R, TB, base, n, g.multi, T)
            live = sample < B
            csample = np.where(live, sample, 0)
            if si == 0 and src is None:
                v = np.zeros((T, NR), dtype=np.complex128)
                for ti in range(T):
                    if not live[ti]:
                        continue
                    E = efac[sample[ti]]
                    pre = 1.0 + 0j
                    for p in range(n):
                        if p not in GP:
                            pre *= E[p, (gbase[ti] >> p) & 1]
                    for j in range(NR):
                        val = pre
                        for r in range(R):
                            val *= E[GP[r], (j >> r) & 1]
                        v[ti, j] = val
            else:
                v = tile[addr].copy()
            for bits, cn in segs:
                for r, ui, ps, _probe, _nc, _qi, _rows in bits:
                    M = utab[ui + (csample if ps else 0)]
                    if M.ndim == 2:
                        M = np.broadcast_to(M, (T, 2, 2))
                    _apply_pair(v, r, M)
                for uniform, a, b in cn:
                    if not uniform:
                        _cnot(v, a, b)
                    else:
                        _swap_where(v, a, ((gbase >> b) & 1).astype(bool))
            if phases:
                v *= _phase(phases, cs, csample, gbase, False)[:, None]
            if si == len(subs) - 1 and measure:
                p = np.abs(v) ** 2
                tot = p.sum(1)
                seglog = (n - R) if g.multi else TB
                for seg in range(T >> seglog):
                    sl = slice(seg << seglog, (seg + 1) << seglog)
                    o = mparts[(t << spt_log) + seg]
                    for r in range(R):
                        o[P[r]] = sum(p[sl, j].sum() for j in range(NR) if not (j >> r) & 1)
                    for i in range(TB):
                        o[G[i]] = tot[sl][((tids[sl] >> i) & 1) == 0].sum()
                    o[K] = tot[sl].sum()
            tile[addr] = v
        state[gi_all[ok]] = tile[swz(ls)][ok]
    return mparts if measure else None


def emulate_adjoint_pass(psi, lam, B, n, R, TB, lpos, words, cs, utab, qtab, grads, seed_w=None):
    """One adjoint pass in place on psi/lam (B*2^n,) complex128; adds per-sample
    Im<lam|Q|psi> sums into ``grads[row][b]`` (the kernel's slot sums, before the 1/2)."""
    g = _Geom(B, n, R, TB, lpos)
    K, T, NR = g.K, g.T, g.NR
    swz = np.vectorize(PL.swz)
    subs = _subpasses(words, R, adjoint=True)
    for t in range(g.ntiles):
        base = g.base(t, n)
        ls = np.arange(1 << K, dtype=np.int64)
        gi_all = base | _dep(ls, g.lpos)
        ok = gi_all < g.total
        src = np.where(ok, psi[np.minimum(gi_all, g.total - 1)], 0)
        if seed_w is not None:
            gi = np.minimum(gi_all, g.total - 1)
            wsum = np.zeros(len(gi))
            for q in range(n):
                wsum += np.where(((gi >> (n - 1 - q)) & 1) == 0, seed_w[gi >> n, q], 0.0)
            lsrc = np.where(ok, 2.0 * wsum * psi[gi], 0)
        else:
            lsrc = np.where(ok, lam[np.minimum(gi_all, g.total - 1)], 0)
        tp = np.zeros(1 << K, complex)
        tl = np.zeros(1 << K, complex)
        tp[swz(ls)] = src
        tl[swz(ls)] = lsrc
        for hdr, phases, segs in subs:
            P, GP, G, gbase, sample, addr, tids = _regs(hdr, R, TB, base, n, g.multi, T)
            live = sample < B
            csample = np.where(live, sample, 0)
            vp, vl = tp[addr].copy(), tl[addr].copy()

            def put(vals, row):
                if row < 0:
                    return
                for ti in range(T):
                    if live[ti]:
                        grads.setdefault(row, np.zeros(B))[sample[ti]] += vals[ti]

            if phases:
                S = np.sum(np.imag(np.conj(vl) * vp), axis=1)
                for pos, ci, fl, row in phases:
                    put(np.where((gbase >> pos) & 1, -S, S), row)
                z = _phase(phases, cs, csample, gbase, True)
                vp *= z[:, None]
                vl *= z[:, None]
            for bits, cn in segs:
                for uniform, a, b in cn:
                    if not uniform:
                        _cnot(vp, a, b)
                        _cnot(vl, a, b)
                    else:
                        cond = ((gbase >> b) & 1).astype(bool)
                        _swap_where(vp, a, cond)
                        _swap_where(vl, a, cond)
                for r, ui, ps, probe, nc, qi, rows in bits:
                    span = B if ps else 1
                    var = csample if ps else np.zeros(T, dtype=np.int64)
                    M = np.zeros((T, 2, 2), complex)
                    for j in range(NR):
                        if (j >> r) & 1:
                            continue
                        j1 = j | (1 << r)
                        pv = np.stack([vp[:, j], vp[:, j1]], 1)
                        lv = np.stack([vl[:, j], vl[:, j1]], 1)
                        M += pv[:, :, None] * np.conj(lv)[:, None, :]
                    for k in range(nc):
                        Q = qtab[qi + k * span + var]
                        put(np.imag(np.einsum("tab,tba->t", Q, M)), rows[k])
                    if not probe:
                        U = utab[ui + var]
                        Ud = np.conj(np.swapaxes(U, -1, -2))
                        _apply_pair(vp, r, Ud)
                        _apply_pair(vl, r, Ud)
            tp[addr] = vp
            tl[addr] = vl
        psi[gi_all[ok]] = tp[swz(ls)][ok]
        lam[gi_all[ok]] = tl[swz(ls)][ok]


def measure_from_partials(mp, B, n, K, L):
    """P(qubit q = 0) from a measuring pass's partials (what the finish kernel does)."""
    probs = np.zeros((B, n))
    if n < K:
        for b in range(B):
            t, s = b >> (K - n), b & ((1 << (K - n)) - 1)
            for q in range(n):
                probs[b, q] = mp[(t << (K - n)) + s, n - 1 - q]
        return probs
    npos = [x for x in range(n) if x not in L]
    loc = {x: j for j, x in enumerate(L)}
    ntl = 1 << len(npos)
    for b in range(B):
        for q in range(n):
            pos = n - 1 - q
            acc = 0.0
            for tl in range(ntl):
                row = mp[b * ntl + tl]
                if pos in loc:
                    acc += row[loc[pos]]
                elif not (tl >> npos.index(pos)) & 1:
                    acc += row[K]
            probs[b, q] = acc
    return probs


def run_forward_words(passes, B, n, R, TB, cs, utab, efac, src=None, measure=True):
    """Forward over encoded passes [(words, lpos)]; returns (state (B, 2^n), probs or None)."""
    state = np.zeros(B << n, dtype=np.complex128)
    cur = None if src is None else src.reshape(-1).astype(np.complex128)
    probs = None
    for i, (words, L) in enumerate(passes):
        last = i == len(passes) - 1
        mp = emulate_pass(state, cur, B, n, R, TB, L, words, cs, utab, efac, measure and last)
        cur = state
        if mp is not None:
            probs = measure_from_partials(mp, B, n, R + TB, L)
    return state.reshape(B, 1 << n), probs


def run_adjoint_words(passes, B, n, R, TB, cs, utab, qtab, psi_final, w):
    """Reverse sweep over encoded adjoint passes [(words, lpos)], the first seeded
    from w = dL/dP; returns ({row: per-sample 1/2 Im<lam|P|psi>}, psi_0, lam_0)."""
    psi = psi_final.reshape(-1).astype(np.complex128).copy()
    lam = np.zeros_like(psi)
    grads = {}
    for k, (wd, L) in enumerate(passes):
        emulate_adjoint_pass(psi, lam, B, n, R, TB, L, wd, cs, utab, qtab, grads, seed_w=w if k == 0 else None)
    return {r: 0.5 * v for r, v in grads.items()}, psi, lam


def check_words(words, R, TB, n, B, ncs, nu, nq, nrows, adjoint, nsub=None):
    """Walks one pass's words with the kernel's cursor and asserts that every
    index the kernel derives from them stays inside its table or register file."""
    K = R + TB
    subs = _subpasses(words, R, adjoint)  # consumes exactly every word or raises
    assert nsub is None or len(subs) == nsub
    posmax = max(n, K)  # a multi-sample tile addresses sample bits above n - 1
    for hdr, phases, segs in subs:
        P = [int(hdr[PL.HDR_P + r]) for r in range(R)]
        G = [int(hdr[PL.HDR_G + i]) for i in range(TB)]
        assert sorted(P + G) == list(range(K)), "register and thread bits must partition the tile"
        for r in range(R):
            assert 0 <= hdr[PL.HDR_GP + r] < posmax
        for i in range(TB):
            assert 0 <= hdr[PL.HDR_GG + i] < posmax
        for pos, ci, fl, row in phases:
            assert 0 <= pos < n
            if not fl & PL.SEG_NOUNAPPLY:
                assert 0 <= ci and ci + ((B - 1) if fl & PL.SEG_PER_SAMPLE else 0) < ncs
            assert row == -1 or 0 <= row < nrows
        for bits, cn in segs:
            for r, ui, ps, probe, nc, qi, rows in bits:
                span = B if ps else 1
                assert 0 <= r < R
                assert 0 <= ui and ui + span <= nu
                if adjoint:
                    assert nc >= 1 and 0 <= qi and qi + nc * span <= nq
                    assert all(-1 <= x < nrows for x in rows)
            for uniform, a, b in cn:
                assert 0 <= a < R
                assert (0 <= b < n) if uniform else (0 <= b < R and b != a)


def check_recipes(recipes, offsets, ncs, nu, nq, generators):
    """Every recipe k_compose reads stays inside cs, utab and (with generators) qtab."""
    for off in offsets:
        u, span, m, q = (int(x) for x in recipes[off: off + 4])
        assert span >= 1 and m >= 1 and 0 <= u and u + span <= nu
        if generators:
            assert 0 <= q and q + m * span <= nq
        for k in range(m):
            kind, ci, cps = (int(x) for x in recipes[off + 4 + 3 * k: off + 7 + 3 * k])
            assert kind in (0, 1, 2, 8, 9, 10)
            if kind < 8:
                assert 0 <= ci and ci + ((span - 1) if cps else 0) < ncs


def emulate_program(prog_gates, n, B, R, TB, cs, efac, src=None, measure=True, cm=None):
    """Plan + encode + compose + emulate the forward; returns (state, probs or None, plan)."""
    pl = PL.plan(n, prog_gates, R, TB, cm)
    tb = PL.TableBuilder(B)
    words = [PL.encode_pass(p, prog_gates, R, TB, lambda s: s, lambda s: False, tb) for p in pl.passes]
    utab, _ = compose(tb.recipes(), B, cs, tb.nu, tb.nq)
    state, probs = run_forward_words(list(zip(words, [p.L for p in pl.passes])), B, n, R, TB, cs, utab, efac,
                                     src=src, measure=measure)
    return state, probs, pl


def emulate_adjoint(prog_gates, n, B, R, TB, cs, psi_final, w, pl, embed_bits=None, embed_rot=0):
    """Reverse sweep of plan ``pl`` from psi_final with co-state seeded by w = dL/dP;
    returns ({row: per-sample 1/2 Im<lam|P|psi>}, psi_0, lam_0).  Rows: rotation
    gates in program order, then the embedded bits."""
    rows = {}
    for gi, gt in enumerate(prog_gates):
        if gt.kind != PL.CNOT:
            rows[gi] = len(rows)
    nrot = len(rows)
    tb = PL.TableBuilder(B)
    words = [PL.encode_pass(p, prog_gates, R, TB, lambda s: s, lambda s: False, tb, reverse=True,
                               row_of=rows.__getitem__) for p in reversed(pl.passes)]
    lpos = [p.L for p in reversed(pl.passes)]
    if embed_bits:
        ggates = [PL.Gate(PL.PGRAD + embed_rot, b, -1, j) for j, b in enumerate(embed_bits)]
        gplan = PL.plan(n, ggates, R, TB)
        for p in gplan.passes:
            words.append(PL.encode_pass(p, ggates, R, TB, lambda s: 0, lambda s: False, tb, reverse=True,
                                           row_of=lambda gi: nrot + gi))
            lpos.append(p.L)
    utab, qtab = compose(tb.recipes(), B, cs, tb.nu, tb.nq)
    return run_adjoint_words(list(zip(words, lpos)), B, n, R, TB, cs, utab, qtab, psi_final, w)
