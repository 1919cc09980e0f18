"""numpy emulation of the fused pass kernel's word semantics (test infrastructure).

Decodes exactly the int32 words the planner sends to ``qs_fused_pass`` and
executes them the way the kernel does (tile gather through the swizzled LDS
index, per-thread register groups, uniform-condition ops, product-state
generation, measurement partials), vectorised over the threads of a tile.
It lets the planner and the encoder be checked against the CPU oracle without
a GPU.
"""
import numpy as np

from qandle_amd._engine import planner as PL


def _dep(vals, positions):
    out = np.zeros_like(vals, dtype=np.int64)
    for i, p in enumerate(positions):
        out |= ((vals >> i) & 1) << p
    return out


def emulate_pass(state, src, B, n, R, TB, lpos, words, cs, efac, measure):
    """state/src: (B*2^n,) complex128 (src None -> generate from efac (B, n, 2)).
    Returns the measurement partials (ntiles*spt, K+1) when ``measure``."""
    K = R + TB
    T, NR = 1 << TB, 1 << R
    multi = n < K
    total = B << n
    if multi:
        lpos = list(range(K))
        ntiles = -(-B // (1 << (K - n)))
        npos = []
    else:
        ntiles = B << (n - K)
        npos = [p for p in range(n) if p not in lpos]
    nn = len(npos)
    spt_log = (K - n) if multi else 0
    mparts = np.zeros((ntiles << spt_log, K + 1))
    tids = np.arange(T, dtype=np.int64)
    swz = np.vectorize(PL.swz)
    for t in range(ntiles):
        if multi:
            base = t << K
        else:
            tl = t & ((1 << nn) - 1)
            base = (t >> nn) << n
            for i in range(nn):
                if (tl >> i) & 1:
                    base |= 1 << npos[i]
        tile = np.zeros(1 << K, dtype=np.complex128)
        if src is not None:
            ls = np.arange(1 << K, dtype=np.int64)
            gi = base | _dep(ls, lpos)
            vals = np.where(gi < total, src[np.minimum(gi, total - 1)], 0)
            tile[swz(ls)] = vals
        w = 0
        nsub = 0
        # count sub-passes from the stream
        pos = 0
        subs = []
        while pos < len(words):
            nops = int(words[pos])
            subs.append(pos)
            pos += PL.HDR + 4 * nops
        for s, w in enumerate(subs):
            hdr = words[w: w + PL.HDR]
            nops = int(hdr[0])
            P = [int(hdr[PL.HDR_P + r]) for r in range(R)]
            SW = [int(hdr[PL.HDR_SWZ + r]) for r in range(R)]
            GP = [int(hdr[PL.HDR_GP + r]) for r in range(R)]
            G = [int(hdr[PL.HDR_G + i]) for i in range(TB)]
            GG = [int(hdr[PL.HDR_GG + i]) for i in range(TB)]
            gl = _dep(tids, G)
            gbase = base | _dep(tids, GG)
            sample = (gbase >> n) if multi else np.full(T, base >> n)
            live = sample < B
            addr = np.zeros((T, NR), dtype=np.int64)
            addr[:, 0] = swz(gl)
            for j in range(1, NR):
                addr[:, j] = addr[:, j & (j - 1)] ^ SW[(j & -j).bit_length() - 1]
            if s == 0 and src is None:
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
            ops = words[w + PL.HDR: w + PL.HDR + 4 * nops].reshape(-1, 4)
            ssafe = np.where(live, sample, 0)
            for code, a, b, coef in ops:
                code, a, b, coef = int(code), int(a), int(b), int(coef)
                kind = code & 0xFF
                if kind == PL.OP_CNOT:
                    for j in range(NR):
                        if (j >> b) & 1 and not (j >> a) & 1:
                            v[:, [j, j | (1 << a)]] = v[:, [j | (1 << a), j]]
                    continue
                if kind == PL.OP_CNOT_UNIFORM:
                    cond = ((gbase >> b) & 1).astype(bool)
                    for j in range(NR):
                        if not (j >> a) & 1:
                            tmp = v[cond, j].copy()
                            v[cond, j] = v[cond, j | (1 << a)]
                            v[cond, j | (1 << a)] = tmp
                    continue
                idx = coef + (ssafe if code & PL.OP_PER_SAMPLE else 0)
                c = cs[idx, 0] * np.ones(T)
                sn = cs[idx, 1] * np.ones(T)
                if kind == PL.OP_RZ_UNIFORM:
                    sgn = np.where((gbase >> a) & 1, -1.0, 1.0)
                    v *= (c - 1j * sgn * sn)[:, None]
                    continue
                for j in range(NR):
                    if (j >> a) & 1:
                        continue
                    j1 = j | (1 << a)
                    a0, a1 = v[:, j].copy(), v[:, j1].copy()
                    if kind == PL.OP_RX:
                        v[:, j] = c * a0 - 1j * sn * a1
                        v[:, j1] = -1j * sn * a0 + c * a1
                    elif kind == PL.OP_RY:
                        v[:, j] = c * a0 - sn * a1
                        v[:, j1] = sn * a0 + c * a1
                    else:
                        v[:, j] = (c - 1j * sn) * a0
                        v[:, j1] = (c + 1j * sn) * a1
            if s == len(subs) - 1 and measure:
                p = np.abs(v) ** 2
                tot = p.sum(1)
                seglog = (n - R) if multi else TB
                for seg in range(T >> seglog):
                    sl = slice(seg << seglog, (seg + 1) << seglog)
                    o = mparts[(t << spt_log) + seg]
                    for r in range(R):
                        o[P[r]] = sum(p[sl, j].sum() for j in range(NR) if not (j >> r) & 1)
                    for i in range(TB):
                        o[G[i]] = tot[sl][((tids[sl] >> i) & 1) == 0].sum()
                    o[K] = tot[sl].sum()
            tile[addr] = v
        ls = np.arange(1 << K, dtype=np.int64)
        gi = base | _dep(ls, lpos)
        ok = gi < total
        state[gi[ok]] = tile[swz(ls)][ok]
    return mparts if measure else None


def emulate_program(prog_gates, n, B, R, TB, cs, efac, src=None, measure=True, cm=None):
    """Plan + encode + emulate; returns (state (B, 2^n), probs (B, n) or None, plan)."""
    from qandle_amd._engine import planner

    pl = planner.plan(n, prog_gates, R, TB, cm)
    state = np.zeros(B << n, dtype=np.complex128)
    cur = None if src is None else src.reshape(-1).astype(np.complex128)
    K = R + TB
    probs = None
    for i, p in enumerate(pl.passes):
        words = planner.encode_pass(p, prog_gates, R, TB, lambda s: s[0] if isinstance(s, tuple) else s,
                                    lambda s: False)
        last = i == len(pl.passes) - 1
        mp = emulate_pass(state, cur, B, n, R, TB, p.L, words, cs, efac, measure and last)
        cur = state
        if mp is not None:
            probs = np.zeros((B, n))
            multi = n < K
            if multi:
                for b in range(B):
                    t, s = b >> (K - n), b & ((1 << (K - n)) - 1)
                    for q in range(n):
                        probs[b, q] = mp[(t << (K - n)) + s, n - 1 - q]
            else:
                npos = [x for x in range(n) if x not in p.L]
                loc = {x: j for j, x in enumerate(p.L)}
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
    return state.reshape(B, 1 << n), probs, pl
