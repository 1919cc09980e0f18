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


# ----------------------------------------------------------------------------
# remaining operators (qs_gates.hip).  A control mask is a set of bit
# positions that must all be 1 for the gate to act (0 = uncontrolled).
# ----------------------------------------------------------------------------

def rot_unitaries(cs: torch.Tensor, kind: int) -> torch.Tensor:
    """(..., 8) float32 column-convention R_kind from (..., 2) (cos, sin) of phi/2."""
    u = torch.empty(cs.shape[:-1] + (8,), dtype=torch.float32, device=cs.device)
    count = cs.numel() // 2
    if count:
        L.check(L.load().qs_rot_unitaries(L.ptr(u), L.ptr(cs), count, kind, L.stream_of(cs)), "qs_rot_unitaries")
    return u


def _unitary(out, src, u, n: int, bit: int, cmask: int, dagger: int):
    B = src.shape[0]
    per_sample = u.dim() == 2
    if per_sample and u.shape[0] != B:
        raise RuntimeError(f"per-sample unitaries: {u.shape[0]} for a batch of {B}")
    L.check(
        L.load().qs_unitary_1q(L.ptr(out), L.ptr(src), B, n, bit, L.ptr(u), 1 if per_sample else 0, cmask, dagger,
                               L.stream_of(src)),
        "qs_unitary_1q",
    )
    return out


class UnitaryFn(torch.autograd.Function):
    """psi' = U psi on one bit position where the control mask holds; U is a
    constant (8,) or per-sample (B, 8) float32 tensor (no gradient)."""

    @staticmethod
    def forward(ctx, state, u, n: int, bit: int, cmask: int):
        state = _c64(state)
        u = u.detach().to(device=state.device, dtype=torch.float32).contiguous()
        ctx.save_for_backward(u)
        ctx.meta = (n, bit, cmask)
        return _unitary(torch.empty_like(state), state, u, n, bit, cmask, 0)

    @staticmethod
    def backward(ctx, g):
        (u,) = ctx.saved_tensors
        n, bit, cmask = ctx.meta
        g = _c64(g)
        return _unitary(torch.empty_like(g), g, u, n, bit, cmask, 1), None, None, None, None


class ControlledRotFn(torch.autograd.Function):
    """Controlled R_P(phi): psi' = R_P(phi) psi on one bit where the control
    mask holds.  dL/dphi = 1/2 Im <g | P (x) |1..1><1..1| | psi'>."""

    @staticmethod
    def forward(ctx, state, phi, n: int, bit: int, kind: int, cmask: int):
        state = _c64(state)
        B = state.shape[0]
        if phi.dim() == 1 and phi.shape[0] != B:
            raise RuntimeError(f"per-sample angle has {phi.shape[0]} entries, state batch is {B}")
        u = rot_unitaries(half_angle_coeffs(phi), kind)
        out = _unitary(torch.empty_like(state), state, u, n, bit, cmask, 0)
        ctx.save_for_backward(out, u)
        ctx.meta = (n, bit, kind, cmask, phi.dim() == 1)
        return out

    @staticmethod
    def backward(ctx, g):
        out, u = ctx.saved_tensors
        n, bit, kind, cmask, per_sample = ctx.meta
        g = _c64(g)
        gin = gphi = None
        if ctx.needs_input_grad[0]:
            gin = _unitary(torch.empty_like(g), g, u, n, bit, cmask, 1)
        if ctx.needs_input_grad[1]:
            B = out.shape[0]
            lib = L.load()
            wb = lib.qs_pauli_im_work_bytes(B, n, 1)
            work = torch.empty(max(wb, 4), dtype=torch.uint8, device=out.device)
            im = torch.empty(B, 1, dtype=torch.float32, device=out.device)
            L.check(
                lib.qs_pauli_im_controlled(L.ptr(im), L.ptr(g), L.ptr(out), B, n, L.i32_array([bit]), 1, kind, cmask,
                                           L.ptr(work), wb, L.stream_of(out)),
                "qs_pauli_im_controlled",
            )
            im = im[:, 0] * 0.5
            gphi = im if per_sample else im.sum()
        return gin, gphi, None, None, None, None


class McxFn(torch.autograd.Function):
    """X on bit ``tbit`` where the control mask holds (an involution)."""

    @staticmethod
    def forward(ctx, state, n: int, tbit: int, cmask: int):
        ctx.meta = (n, tbit, cmask)
        return McxFn._run(_c64(state), n, tbit, cmask)

    @staticmethod
    def _run(x, n, tbit, cmask):
        out = torch.empty_like(x)
        L.check(L.load().qs_mcx(L.ptr(out), L.ptr(x), x.shape[0], n, tbit, cmask, L.stream_of(x)), "qs_mcx")
        return out

    @staticmethod
    def backward(ctx, g):
        n, tbit, cmask = ctx.meta
        return McxFn._run(_c64(g), n, tbit, cmask), None, None, None


class PhaseFlipFn(torch.autograd.Function):
    """Negates the amplitudes whose bits ``mask`` are all 1 (diagonal, real)."""

    @staticmethod
    def forward(ctx, state, n: int, mask: int):
        ctx.meta = (n, mask)
        return PhaseFlipFn._run(_c64(state), n, mask)

    @staticmethod
    def _run(x, n, mask):
        out = torch.empty_like(x)
        L.check(L.load().qs_phase_flip(L.ptr(out), L.ptr(x), x.shape[0], n, mask, L.stream_of(x)), "qs_phase_flip")
        return out

    @staticmethod
    def backward(ctx, g):
        n, mask = ctx.meta
        return PhaseFlipFn._run(_c64(g), n, mask), None, None


class SwapFn(torch.autograd.Function):
    """Exchanges bits ``abit`` and ``bbit`` where the control mask holds (an involution)."""

    @staticmethod
    def forward(ctx, state, n: int, abit: int, bbit: int, cmask: int):
        ctx.meta = (n, abit, bbit, cmask)
        return SwapFn._run(_c64(state), n, abit, bbit, cmask)

    @staticmethod
    def _run(x, n, abit, bbit, cmask):
        out = torch.empty_like(x)
        L.check(L.load().qs_swap(L.ptr(out), L.ptr(x), x.shape[0], n, abit, bbit, cmask, L.stream_of(x)), "qs_swap")
        return out

    @staticmethod
    def backward(ctx, g):
        n, abit, bbit, cmask = ctx.meta
        return SwapFn._run(_c64(g), n, abit, bbit, cmask), None, None, None, None


class ResetFn(torch.autograd.Function):
    """Reset of one bit position (qs_reset); not unitary, gradient by qs_reset_backward."""

    @staticmethod
    def forward(ctx, state, n: int, bit: int):
        state = _c64(state)
        out = torch.empty_like(state)
        L.check(L.load().qs_reset(L.ptr(out), L.ptr(state), state.shape[0], n, bit, L.stream_of(state)), "qs_reset")
        ctx.save_for_backward(state)
        ctx.meta = (n, bit)
        return out

    @staticmethod
    def backward(ctx, g):
        (state,) = ctx.saved_tensors
        n, bit = ctx.meta
        g = _c64(g)
        gin = torch.empty_like(g)
        L.check(L.load().qs_reset_backward(L.ptr(gin), L.ptr(g), L.ptr(state), g.shape[0], n, bit, L.stream_of(g)),
                "qs_reset_backward")
        return gin, None, None


class AmplitudeEmbedFn(torch.autograd.Function):
    """(B, d) float32 features -> (B, 2^n) real amplitudes: padded with ``pad``
    to 2^n (or cropped), L2-normalised when ``normalize``."""

    @staticmethod
    def forward(ctx, x, n: int, pad: float, normalize: bool):
        x = x.detach().to(torch.float32).contiguous()
        B, d = x.shape
        out = torch.empty(B, 2 ** n, dtype=torch.complex64, device=x.device)
        L.check(L.load().qs_amplitude_embed(L.ptr(out), L.ptr(x), B, d, n, d, float(pad), int(normalize),
                                            L.stream_of(x)), "qs_amplitude_embed")
        ctx.save_for_backward(x)
        ctx.meta = (n, float(pad), int(normalize))
        return out

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        n, pad, normalize = ctx.meta
        g = _c64(g)
        gx = torch.empty_like(x)
        B, d = x.shape
        L.check(L.load().qs_amplitude_embed_backward(L.ptr(gx), L.ptr(g), L.ptr(x), B, d, n, d, pad, normalize,
                                                     L.stream_of(x)), "qs_amplitude_embed_backward")
        return gx, None, None, None


class Abs2Fn(torch.autograd.Function):
    """|psi|^2 elementwise (float32), gradient 2 g psi."""

    @staticmethod
    def forward(ctx, state):
        state = _c64(state)
        out = torch.empty(state.shape, dtype=torch.float32, device=state.device)
        L.check(L.load().qs_abs2(L.ptr(out), L.ptr(state), state.numel(), L.stream_of(state)), "qs_abs2")
        ctx.save_for_backward(state)
        return out

    @staticmethod
    def backward(ctx, g):
        (state,) = ctx.saved_tensors
        g = g.to(torch.float32).contiguous()
        gin = torch.empty_like(state)
        L.check(L.load().qs_abs2_backward(L.ptr(gin), L.ptr(g), L.ptr(state), state.numel(), L.stream_of(state)),
                "qs_abs2_backward")
        return gin
