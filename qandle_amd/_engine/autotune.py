"""Autotuner for the fused pass engine.

Picks, per (program structure, qubit count, batch), the kernel shape (register
bits R, thread bits TB) and the planner's cost model (VALU budget per pass,
number of always-local low positions) by timing every candidate on the
device with HIP events and keeping the fastest.  Results live in memory and,
when ``config.AUTOTUNE_CACHE`` names a file, in a JSON cache keyed by the
device name, so a tuned shape survives restarts.
"""
from __future__ import annotations

import dataclasses
import hashlib
import json
import os
import threading

import torch

from .. import config
from . import planner as PL

_lock = threading.Lock()
_memo = {}


def candidates(n: int):
    """(R, TB, budget, min_low) candidates for width n.

    With specialised kernels every candidate is a compilation, so the list is
    short and spans the fusion budgets those kernels favour; the interpreting
    kernels are tuned over lighter budgets and the always-local low bits."""
    out = []
    shapes = [(5, 6), (4, 8), (5, 8), (4, 7)] if n >= 5 else [(n, 8)]
    if getattr(config, "SPECIALISE", False):
        for R, TB in shapes[:3]:
            if R <= n:
                out += [(R, TB, 256.0, 4), (R, TB, 1024.0, 4)]
        return out
    for R, TB in shapes:
        if R > n:
            continue
        for budget in (64.0, 112.0, 160.0, 240.0):
            for low in (4, 6):
                out.append((R, TB, budget, low))
    return out


def fingerprint(prog) -> str:
    h = hashlib.sha1()
    h.update(repr((prog.n, prog.init[0], prog.measure)).encode())
    for g in prog.gates:
        h.update(b"%d,%d,%d;" % (g.kind, g.pos, g.ctrl))
    return h.hexdigest()[:16]


def _cache_path():
    return getattr(config, "AUTOTUNE_CACHE", None)


def _load_disk():
    path = _cache_path()
    if not path or not os.path.exists(path):
        return {}
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_disk(key, value):
    path = _cache_path()
    if not path:
        return
    data = _load_disk()
    data[key] = value
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def key_for(prog, B, device) -> str:
    name = torch.cuda.get_device_name(device).replace(" ", "_")
    return f"{name}|{fingerprint(prog)}|n{prog.n}|B{B}"


def lookup(prog, B, device):
    k = key_for(prog, B, device)
    with _lock:
        if k in _memo:
            return _memo[k]
        disk = _load_disk().get(k)
        if disk:
            val = (disk["R"], disk["TB"], PL.CostModel(budget=disk["budget"], min_low=disk["min_low"]))
            _memo[k] = val
            return val
    return None


def tune(prog, inp, reps: int = 3, cands=None):
    """Time every candidate on ``inp`` (forward only) and remember the fastest."""
    from . import fused

    cands = cands or candidates(prog.n)
    cs = fused.build_coeffs(prog, inp)
    efac = fused.product_factors(prog, inp) if (prog.init[0] == "embed" or inp.state is None) else None
    best, best_ms, table = None, float("inf"), []
    stream = torch.cuda.current_stream(inp.device)
    for R, TB, budget, low in cands:
        cm = PL.CostModel(budget=budget, min_low=low)
        fused.execute(prog, inp, cs, efac, R, TB, cm=cm, want_state=False)  # plan + warm-up
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fused.execute(prog, inp, cs, efac, R, TB, cm=cm, want_state=False)
        e1.record(stream)
        torch.cuda.synchronize(inp.device)
        ms = e0.elapsed_time(e1) / reps
        table.append({"R": R, "TB": TB, "budget": budget, "min_low": low, "ms": ms})
        if ms < best_ms:
            best, best_ms = (R, TB, cm), ms
    k = key_for(prog, inp.B, inp.device)
    with _lock:
        _memo[k] = best
    _store_disk(k, {"R": best[0], "TB": best[1], "budget": best[2].budget, "min_low": best[2].min_low,
                    "ms": best_ms})
    return best, table


def tune_uneven(make_batch, cands=((4, 8, 1024.0), (4, 8, 2048.0), (4, 7, 1024.0), (5, 8, 1024.0), (5, 6, 1024.0)), reps=3,
                backward=True):
    """Build an uneven batch per candidate (via ``make_batch(R, TB, cost)``), time a
    captured replay of each, return (best batch, table)."""
    best, best_ms, table = None, float("inf"), []
    for R, TB, budget in cands:
        try:
            ub = make_batch(R, TB, PL.CostModel(budget=budget, min_low=4))
        except ValueError:
            continue  # a circuit narrower than R register bits
        ub.capture(backward=backward)
        ub.replay(backward=backward)
        dev = ub.device
        stream = torch.cuda.current_stream(dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize(dev)
        e0.record(stream)
        for _ in range(reps):
            ub.replay(backward=backward)
        e1.record(stream)
        torch.cuda.synchronize(dev)
        ms = e0.elapsed_time(e1) / reps
        table.append({"R": R, "TB": TB, "budget": budget, "ms": ms})
        if ms < best_ms:
            best, best_ms = ub, ms
        else:
            del ub
    return best, table
