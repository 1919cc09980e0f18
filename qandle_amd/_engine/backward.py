"""Adjoint backward of a fused program, behind one torch.autograd.Function.

Forward: the fused passes, keeping the final state and a copy of the state
every ``checkpoint_gates`` gates (at pass boundaries).

Backward (reverse sweep over the same passes, ``qs_adjoint_pass``):

* co-state seed: lambda = 2 psi * sum_q w_q [bit n-1-q = 0] for a
  MeasureProbability output (w = dL/dP), or dL/dpsi for a state output;
  when the compiled kernels fold trailing CNOTs into the readout, psi is the
  state before them and the measured bit is the parity of its readout mask;
* per pass, in reverse: for every gate, 1/2 Im<lambda|P|psi> at its output,
  then psi <- U^dag psi, lambda <- U^dag lambda;
* where a checkpoint exists, psi is reset to it, which bounds the float32
  drift of un-applying gates on a deep circuit;
* AngleEmbedding inputs: dL/dx_j = 1/2 Im<lambda_0|P_j|psi_0> (the embedded
  factors commute, so every column comes from the one final pair);
* an initial state that needs a gradient gets lambda_0.

Gradients of remapped parameters flow through torch (``remap(theta)`` is
computed outside the Function).
"""
from __future__ import annotations

import torch

from .. import config
from . import fused
from . import lib as L


# When set to a list, the backward sweep appends (start, end, algorithmic bytes)
# for every adjoint kernel launch: HIP events recorded on the launch stream.
# Measurement only (the benchmark's roofline of the adjoint kernels).
launch_events = None


def checkpoint_after(prog, pl, every_gates: int):
    """Pass indices after which the forward state is stored."""
    out = set()
    done = 0
    next_mark = every_gates
    for i, p in enumerate(pl.passes[:-1]):
        done += len(p.gates)
        if every_gates > 0 and done >= next_mark:
            out.add(i)
            while next_mark <= done:
                next_mark += every_gates
    return out


def default_checkpoint_gates(n: int) -> int:
    """CHECKPOINT_LAYERS layers of (RX, RY, RZ per qubit + a CNOT per qubit)."""
    return int(config.CHECKPOINT_LAYERS) * 4 * n


class FusedCircuitFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, prog, R, TB, cm, B, batched, names, phi_param, embed_x, state, *named_vals):
        named = dict(zip(names, [v.detach() for v in named_vals]))
        inp = fused.Inputs(B, batched, phi_param.device, phi_param.detach(), named,
                           None if embed_x is None else embed_x.detach(),
                           None if state is None else state.detach())
        cs = fused.build_coeffs(prog, inp)
        efac = fused.product_factors(prog, inp) if (prog.init[0] == "embed" or state is None) else None
        jit_on = fused.use_jit(prog.n, R, TB)
        fold = fused.folds(prog, jit_on)
        pl = fused.get_plan(prog, R, TB, cm, fold)
        every = config.CHECKPOINT_GATES if getattr(config, "CHECKPOINT_GATES", None) else default_checkpoint_gates(prog.n)
        marks = checkpoint_after(prog, pl, every)
        ckpts = {}

        def keep(i, psi):
            if i in marks:  # the pass's output stays as the checkpoint; later passes write a new buffer
                ckpts[i] = psi
                return torch.empty_like(psi)
            return None

        psi, probs = fused.execute(prog, inp, cs, efac, R, TB, cm=cm, want_state=True, checkpoints=keep,
                                   specialise=jit_on, fold=fold)
        ctx.prog, ctx.R, ctx.TB, ctx.cm, ctx.B, ctx.names = prog, R, TB, cm, B, names
        ctx.jit, ctx.fold = jit_on, fold
        ctx.ck_index = sorted(ckpts)
        ctx.named_shapes = [v.shape for v in named_vals]
        ctx.embed_shape = None if embed_x is None else embed_x.shape
        ctx.state_shape = None if state is None else state.shape
        ctx.save_for_backward(psi, cs, *[ckpts[i] for i in ctx.ck_index])
        if prog.measure:
            return probs
        return psi

    @staticmethod
    def backward(ctx, gout):
        prog, R, TB, B = ctx.prog, ctx.R, ctx.TB, ctx.B
        saved = ctx.saved_tensors
        psi_final, cs = saved[0], saved[1]
        ckpts = dict(zip(ctx.ck_index, saved[2:]))
        n = prog.n
        lib = L.load()
        st = L.stream_of(psi_final)
        # the compiled steps read psi from psi_src and write psi: the first reads the saved final
        # state and each reset reads its checkpoint, with no copies; the interpreting steps work in place
        psi = torch.empty_like(psi_final) if ctx.jit else psi_final.clone()
        psi_src = psi_final
        lam = torch.empty_like(psi)
        seed = None
        if prog.measure:
            # the first adjoint pass builds lambda from psi and w = dL/dP while loading
            seed = gout.to(torch.float32).contiguous()
        else:
            lam.copy_(gout.to(torch.complex64).resolve_conj())
        pl = fused.get_plan(prog, R, TB, ctx.cm, ctx.fold)
        rows = fused.rotation_rows(prog)
        nrows = len(rows)
        embed = prog.init[0] == "embed" and ctx.needs_input_grad[8]
        ws, nmain = fused.adjoint_words(prog, pl, B, psi.device, embed, ctx.cm)
        embed_rows = len(prog.init[2]) if embed else 0
        total_rows = nrows + embed_rows
        utab, qtab = fused.compose_tables(ws, cs)
        wb = lib.qs_adjoint_work_bytes(B, n, R, TB, total_rows)
        gpart = torch.empty(max(wb // 4, 1), dtype=torch.float32, device=psi.device)
        npass = len(pl.passes)
        kern = None
        ltab = None
        if ctx.jit:
            kern = specialised_adjoint(prog, pl, ws, B, embed, seed is not None, ctx.fold)
            ntiles, NS = adjoint_geometry(B, n, R, TB)
            idx = kern.lane_index_on(psi.device)
            if idx is not None:  # U^dagger operands of the matrix-core unitaries, step by step
                ltab = torch.empty(idx.numel() * 16, dtype=torch.float32, device=psi.device)
                L.check(lib.qs_lane_coeffs(L.ptr(ltab), L.ptr(utab), L.ptr(idx), idx.numel(), 1, st),
                        "qs_lane_coeffs")
        events = launch_events
        for k, (w, nsub, lpos) in enumerate(ws.passes):
            seed_ptr = None if (k > 0 or seed is None) else L.ptr(seed)
            if events is not None:
                ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                ev[0].record(torch.cuda.current_stream(psi.device))
            if kern is not None:
                kern.launch(k, L.ptr(psi), L.ptr(lam), L.ptr(utab), L.ptr(qtab), L.ptr(cs), L.ptr(gpart), seed_ptr,
                            B, NS, ntiles, st, None if ltab is None else L.ptr(ltab), L.ptr(psi_src))
                psi_src = psi
            else:
                L.check(
                    lib.qs_adjoint_pass(
                        L.ptr(psi), L.ptr(lam), B, n, R, TB,
                        lpos.ctypes.data_as(L.ctypes.POINTER(L.ctypes.c_int32)), L.ptr(w), nsub,
                        L.ptr(cs), L.ptr(utab), L.ptr(qtab), L.ptr(gpart), total_rows, seed_ptr, st,
                    ),
                    "qs_adjoint_pass",
                )
            if events is not None:
                ev[1].record(torch.cuda.current_stream(psi.device))
                # psi and lambda are read and written once per step; a seeded step reads psi only
                events.append((ev[0], ev[1], B * (2 ** n) * (24 if seed_ptr is not None else 32)))
            i = npass - 1 - k  # forward pass this step undid
            if k < nmain and (i - 1) in ckpts:
                if kern is not None:
                    psi_src = ckpts[i - 1]
                else:
                    psi.copy_(ckpts[i - 1])
        per = torch.empty(max(total_rows, 1), B, dtype=torch.float32, device=psi.device)
        tot = torch.empty(max(total_rows, 1), dtype=torch.float32, device=psi.device)
        L.check(lib.qs_adjoint_finish(L.ptr(per), L.ptr(tot), L.ptr(gpart), total_rows, B, n, R, TB, st),
                "qs_adjoint_finish")

        scatter = gradient_scatter(prog, rows, psi.device)
        # parameters (shared angles): one row each
        g_param = None
        if ctx.needs_input_grad[7]:
            g_param = torch.zeros(len(prog.param_slots), dtype=torch.float32, device=psi.device)
            if scatter["param"] is not None:
                idx, src_rows = scatter["param"]
                g_param.index_add_(0, idx, tot[src_rows])
        # named inputs: per-sample rows summed over the gates that use the name
        g_named = []
        for k, name in enumerate(ctx.names):
            if not ctx.needs_input_grad[10 + k]:
                g_named.append(None)
                continue
            rr = scatter["named"].get(name)
            gv = per[rr].sum(0) if rr is not None else torch.zeros(B, device=psi.device)
            shape = ctx.named_shapes[k]
            g_named.append(gv.sum().reshape(shape) if len(shape) == 0 else gv.reshape(shape))
        g_embed = None
        if embed_rows:
            gx = per[nrows:total_rows].t()  # (B, d), already 1/2 Im<lam|P_j|psi>
            shp = ctx.embed_shape
            g_embed = gx.sum(0) if len(shp) == 1 else gx.reshape(shp)
        g_state = None
        if ctx.needs_input_grad[9] and prog.init[0] != "embed":
            shp = ctx.state_shape
            g_state = lam.sum(0) if len(shp) == 1 else (lam if lam.shape == shp else lam.sum(0, keepdim=True))
        return (None, None, None, None, None, None, None, g_param, g_embed, g_state, *g_named)


def gradient_scatter(prog, rows, device):
    """Index tensors that route gradient rows to angles, uploaded once per program
    and device: ``param`` = (angle index, row) of the unnamed rotations (None
    when there are none), ``named`` = rows of each named input's gates."""
    key = ("grad_scatter", str(device))
    hit = prog.words_cache.get(key)
    if hit is None:
        slot_pos = {s_: i for i, s_ in enumerate(prog.param_slots)}
        param_rows, named = [], {}
        for gi, g in enumerate(prog.gates):
            if g.kind == 4:  # CNOT
                continue
            sl = prog.slots[g.slot]
            if sl.kind == "param":
                param_rows.append((slot_pos[g.slot], rows[gi]))
            elif sl.kind == "named":
                named.setdefault(sl.name, []).append(rows[gi])
        hit = {
            "param": (L.upload([p for p, _ in param_rows], torch.long, device),
                      L.upload([r for _, r in param_rows], torch.long, device)) if param_rows else None,
            "named": {nm: L.upload(rr, torch.long, device) for nm, rr in named.items()},
        }
        prog.words_cache[key] = hit
    return hit


def adjoint_geometry(B: int, n: int, R: int, TB: int):
    """(tiles, reduction slots per gradient row) of an adjoint pass, as the
    library's adjoint kernels and qs_adjoint_finish lay them out."""
    K = R + TB
    multi = n < K
    seglog = (n - R) if multi else TB
    wlog = min(seglog, 6)
    ntiles = ((B + (1 << (K - n)) - 1) >> (K - n)) if multi else (B << (n - K))
    return ntiles, ntiles * (1 << (TB - wlog))


def specialised_adjoint(prog, pl, ws, B, embed, seeded, fold=False):
    """Run-time compiled kernels of this plan's adjoint sweep (cached on the
    program); with ``fold`` the seed reads the readout masks of the folded
    trailing CNOTs (fused.readout_fold)."""
    from . import jit

    opts = getattr(config, "SPECIALISE_ADJOINT_OPTIONS", {}) or {}
    key = ("jit_adj", id(pl), B, embed, seeded, fold, tuple(sorted(opts.items())))
    hit = prog.words_cache.get(key)
    if hit is None:
        masks = fused.readout_fold(prog)[1] if fold else None
        hit = jit.AdjointKernels(ws.host_passes(), prog.n, pl.R, pl.TB, seeded, readout=masks, **opts)
        prog.words_cache[key] = hit
    return hit


def run_with_grad(prog, inp, R, TB, cm=None):
    names = list(inp.named)
    vals = [inp.named[k] for k in names]
    out = FusedCircuitFn.apply(prog, R, TB, cm or fused.GRAD_COST, inp.B, inp.batched, names, inp.phi_param,
                               inp.embed_x, inp.state, *vals)
    if prog.measure:
        return None, out
    return out, None
