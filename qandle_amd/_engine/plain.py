"""Plain per-gate path: one HIP kernel launch (one HBM pass) per operator.

Each operator module of the drop-in API runs through these autograd Functions
when it is called on its own; ``Circuit.forward`` uses the fused program path
instead (see ``fused.py``).  Backward rules (torch's convention: for a real
loss the gradient of a complex tensor is dL/dRe + i dL/dIm):

* rotation psi' = U psi:   grad_psi = U^dag g,
                           dL/dphi = 1/2 Im <g | P | psi'>
* CNOT (self-inverse):     grad_psi = CNOT g
* AngleEmbedding:          dL/dx_j = 1/2 Im <g | P_j | psi>  (the factors commute)
* MeasureProbability:      grad_psi = 2 psi * sum_q w_q [bit n-1-q = 0]
"""
from __future__ import annotations

import torch

from . import lib as L

KIND = {"x": 0, "y": 1, "z": 2}


def half_angle_coeffs(phi: torch.Tensor) -> torch.Tensor:
    """(..., 2) float32 tensor of (cos(phi/2), sin(phi/2)) computed on the device."""
    phi = phi.detach().to(torch.float32).contiguous()
    cs = torch.empty(phi.shape + (2,), dtype=torch.float32, device=phi.device)
    if phi.numel():
        L.check(
            L.load().qs_half_angle_coeffs(L.ptr(cs), L.ptr(phi), phi.numel(), L.stream_of(phi)),
            "qs_half_angle_coeffs",
        )
    return cs


def _c64(state: torch.Tensor) -> torch.Tensor:
    if state.dtype != torch.complex64:
        raise TypeError(f"state must be complex64, got {state.dtype}")
    return state.resolve_conj().contiguous()


def pauli_im(lam: torch.Tensor, psi: torch.Tensor, n: int, bits, pauli: int) -> torch.Tensor:
    """(B, d) float32: Im <lam_b | P_j | psi_b> for each bit position in ``bits``."""
    B = psi.shape[0]
    d = len(bits)
    lib = L.load()
    wb = lib.qs_pauli_im_work_bytes(B, n, d)
    work = torch.empty(max(wb, 4), dtype=torch.uint8, device=psi.device)
    out = torch.empty(B, d, dtype=torch.float32, device=psi.device)
    L.check(
        lib.qs_pauli_im(
            L.ptr(out), L.ptr(lam), L.ptr(psi), B, n, L.i32_array(bits), d, pauli,
            L.ptr(work), wb, L.stream_of(psi),
        ),
        "qs_pauli_im",
    )
    return out


class RotFn(torch.autograd.Function):
    """psi' = R_P(phi) psi on one bit position; phi is () (shared) or (B,)."""

    @staticmethod
    def forward(ctx, state, phi, n: int, bit: int, kind: int):
        state = _c64(state)
        B = state.shape[0]
        per_sample = phi.dim() == 1
        if per_sample and phi.shape[0] != B:
            raise RuntimeError(f"per-sample angle has {phi.shape[0]} entries, state batch is {B}")
        cs = half_angle_coeffs(phi)
        out = torch.empty_like(state)
        L.check(
            L.load().qs_rot(
                L.ptr(out), L.ptr(state), B, n, bit, kind, L.ptr(cs), 1 if per_sample else 0, 0,
                L.stream_of(state),
            ),
            "qs_rot",
        )
        ctx.save_for_backward(out, cs)
        ctx.meta = (n, bit, kind, per_sample)
        return out

    @staticmethod
    def backward(ctx, g):
        out, cs = ctx.saved_tensors
        n, bit, kind, per_sample = ctx.meta
        g = _c64(g)
        B = g.shape[0]
        gin = gphi = None
        if ctx.needs_input_grad[0]:
            gin = torch.empty_like(g)
            L.check(
                L.load().qs_rot(
                    L.ptr(gin), L.ptr(g), B, n, bit, kind, L.ptr(cs), 1 if per_sample else 0, 1,
                    L.stream_of(g),
                ),
                "qs_rot(inverse)",
            )
        if ctx.needs_input_grad[1]:
            im = pauli_im(g, out, n, [bit], kind)[:, 0] * 0.5
            gphi = im if per_sample else im.sum()
        return gin, gphi, None, None, None


class CnotFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state, n: int, cbit: int, tbit: int):
        state = _c64(state)
        out = torch.empty_like(state)
        L.check(
            L.load().qs_cnot(L.ptr(out), L.ptr(state), state.shape[0], n, cbit, tbit, L.stream_of(state)),
            "qs_cnot",
        )
        ctx.meta = (n, cbit, tbit)
        return out

    @staticmethod
    def backward(ctx, g):
        n, cbit, tbit = ctx.meta
        g = _c64(g)
        gin = torch.empty_like(g)
        L.check(L.load().qs_cnot(L.ptr(gin), L.ptr(g), g.shape[0], n, cbit, tbit, L.stream_of(g)), "qs_cnot")
        return gin, None, None, None


class EmbedFn(torch.autograd.Function):
    """(B, d) angles -> (B, 2^n) product state (x)_j R(x_j)|0>."""

    @staticmethod
    def forward(ctx, x, n: int, bits, rot: int):
        x = x.detach().to(torch.float32).contiguous()
        B = x.shape[0]
        out = torch.empty(B, 2 ** n, dtype=torch.complex64, device=x.device)
        L.check(
            L.load().qs_embed_angle(
                L.ptr(out), L.ptr(x), B, n, L.i32_array(bits), len(bits), rot, x.shape[1],
                L.stream_of(x),
            ),
            "qs_embed_angle",
        )
        ctx.save_for_backward(out)
        ctx.meta = (n, tuple(bits), rot)
        return out

    @staticmethod
    def backward(ctx, g):
        (out,) = ctx.saved_tensors
        n, bits, rot = ctx.meta
        gx = None
        if ctx.needs_input_grad[0]:
            gx = pauli_im(_c64(g), out, n, list(bits), rot) * 0.5
        return gx, None, None, None


class MeasureProbFn(torch.autograd.Function):
    """(B, 2^n) complex64 -> (B, n) float32, P(qubit q = 0)."""

    @staticmethod
    def forward(ctx, state, n: int):
        state = _c64(state)
        B = state.shape[0]
        lib = L.load()
        wb = lib.qs_measure_prob_work_bytes(B, n)
        work = torch.empty(max(wb, 4), dtype=torch.uint8, device=state.device)
        out = torch.empty(B, n, dtype=torch.float32, device=state.device)
        L.check(
            lib.qs_measure_prob(L.ptr(out), L.ptr(state), B, n, L.ptr(work), wb, L.stream_of(state)),
            "qs_measure_prob",
        )
        ctx.save_for_backward(state)
        ctx.n = n
        return out

    @staticmethod
    def backward(ctx, w):
        (state,) = ctx.saved_tensors
        n = ctx.n
        w = w.to(torch.float32).contiguous()
        lam = torch.empty_like(state)
        L.check(
            L.load().qs_measure_prob_backward(
                L.ptr(lam), L.ptr(state), L.ptr(w), n, state.shape[0], n, L.stream_of(state)
            ),
            "qs_measure_prob_backward",
        )
        return lam, None
