"""Forward entry of the fused engine and its autograd Function.

Without gradients the circuit runs as fused passes only.  When any angle,
input or the initial state needs a gradient, the forward also keeps the final
state (and checkpoints), and the backward is the adjoint sweep implemented in
``backward.py``.
"""
from __future__ import annotations

import torch

from .. import config
from . import fused


def _needs_grad(inp) -> bool:
    if not torch.is_grad_enabled():
        return False
    if inp.phi_param.requires_grad:
        return True
    if any(v.requires_grad for v in inp.named.values()):
        return True
    if inp.embed_x is not None and inp.embed_x.requires_grad:
        return True
    return inp.state is not None and inp.state.requires_grad


def forward(prog, inp):
    grad = _needs_grad(inp)
    R, TB, cm = fused.choose_config(prog, inp, grad)
    if not grad:
        cs = fused.build_coeffs(prog, inp)
        efac = fused.product_factors(prog, inp) if (prog.init[0] == "embed" or inp.state is None) else None
        psi, probs = fused.execute(prog, inp, cs, efac, R, TB, cm=cm, want_state=not prog.measure)
        return psi, probs
    from . import backward

    return backward.run_with_grad(prog, inp, R, TB, cm)
