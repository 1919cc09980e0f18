"""GPU parity of the adjoint backward (fused path) against the FP64 oracle.

The reference has no gradient tests (its tests only assert requires_grad);
gradients are pinned by the FP64 adjoint of ``oracle/fp64.py``, which is itself
pinned by central finite differences and parameter shift.  Bar: 1e-4 absolute
for dL/dtheta and dL/dx with L = sum(w * MeasureProbability).
"""
import numpy as np
import pytest
import torch

import qandle_amd as qandle
from oracle import fp64
from qandle_amd import config, workloads
from qandle_amd._engine import fused

pytestmark = pytest.mark.gpu
GTOL = 1e-4


def grads_fused(w, device, layers_path=False):
    circ = qandle.Circuit(w.layers, num_qubits=w.n).to(device)
    x = w.inputs["x"].to(device).requires_grad_(True)
    if layers_path:
        P = circ.circuit.forward_layers(x=x)
    else:
        P = circ(x=x)
    loss = (P * w.loss_w.to(device)).sum()
    loss.backward()
    thetas = [m for m in circ.circuit.layers if isinstance(m, qandle.operators.BuiltParametrizedOperator)]
    return x.grad.cpu().numpy(), np.array([m.theta.grad.item() for m in thetas]), P.detach().cpu().numpy()


def oracle_grads(w):
    ops = w.oracle_ops()
    adj = fp64.adjoint_grads(w.n, ops, w.loss_w.double().numpy())
    gates = np.array([a for a, op in zip(adj[1:], ops[1:]) if op[0] != "cnot"])
    return adj[0], gates


@pytest.mark.parametrize("n,B,depth", [(4, 3, 2), (8, 5, 3), (11, 2, 2), (13, 3, 2), (16, 2, 2)])
def test_adjoint_vs_fp64(device, n, B, depth):
    w = workloads.rxryrz_ladder(n=n, B=B, depth=depth, seed=100 + n, loss=True)
    gx, gth, _ = grads_fused(w, device)
    ox, oth = oracle_grads(w)
    ex, eth = np.abs(gx - ox).max(), np.abs(gth - oth).max()
    print(f"adjoint n={n} B={B} depth={depth}: max|dx err|={ex:.2e} max|dtheta err|={eth:.2e}")
    assert ex < GTOL and eth < GTOL


def test_adjoint_matches_plain_autograd(device):
    w = workloads.embed_ry_ring(n=8, B=6, depth=4, seed=3)
    w.loss_w = torch.randn(w.B, w.n, generator=torch.Generator().manual_seed(1))
    gx_f, gth_f, P_f = grads_fused(w, device)
    gx_p, gth_p, P_p = grads_fused(w, device, layers_path=True)
    assert np.abs(P_f - P_p).max() < 1e-5
    assert np.abs(gx_f - gx_p).max() < GTOL
    assert np.abs(gth_f - gth_p).max() < GTOL


def test_checkpointing_changes_nothing_but_drift(device):
    w = workloads.rxryrz_ladder(n=14, B=3, depth=6, seed=9, loss=True)
    old = config.CHECKPOINT_GATES
    try:
        config.CHECKPOINT_GATES = 10 ** 9  # no checkpoints
        gx0, gth0, _ = grads_fused(w, device)
        config.CHECKPOINT_GATES = 30  # a checkpoint after almost every pass
        gx1, gth1, _ = grads_fused(w, device)
    finally:
        config.CHECKPOINT_GATES = old
    assert np.abs(gx0 - gx1).max() < 1e-5 and np.abs(gth0 - gth1).max() < 1e-5
    ox, oth = oracle_grads(w)
    assert np.abs(gx1 - ox).max() < GTOL and np.abs(gth1 - oth).max() < GTOL


def test_named_per_sample_and_state_gradients(device):
    """Re-uploading (B,) inputs, a 0-d named input and an input state that needs a gradient."""
    gen = torch.Generator().manual_seed(12)
    n, B = 12, 4
    psi0 = torch.randn(B, 2 ** n, dtype=torch.cfloat, generator=gen)
    psi0 = psi0 / torch.linalg.norm(psi0, dim=-1, keepdim=True)
    r = torch.rand(B, generator=gen) * 3
    s0 = torch.rand((), generator=gen) * 3
    th = [float(v) for v in torch.rand(n, generator=gen) * 6]
    layers = [qandle.RY(q, theta=th[q], remapping=None) for q in range(n)]
    layers += [qandle.RX(3, name="r", remapping=None), qandle.CNOT(3, 9), qandle.RZ(9, name="r"),
               qandle.RY(0, name="s"), qandle.CNOT(11, 0), qandle.MeasureProbability()]
    circ = qandle.Circuit(layers, num_qubits=n).to(device)
    wt = torch.randn(B, n, generator=gen)
    st = psi0.to(device).requires_grad_(True)
    rr = r.to(device).requires_grad_(True)
    ss = s0.to(device).requires_grad_(True)
    P = circ(st, r=rr, s=ss)
    (P * wt.to(device)).sum().backward()
    # the same through the plain per-gate path (its autograd is pinned separately)
    st2 = psi0.to(device).requires_grad_(True)
    rr2 = r.to(device).requires_grad_(True)
    ss2 = s0.to(device).requires_grad_(True)
    P2 = circ.circuit.forward_layers(st2, r=rr2, s=ss2)
    (P2 * wt.to(device)).sum().backward()
    assert (P - P2).abs().max().item() < 1e-5
    assert (rr.grad - rr2.grad).abs().max().item() < GTOL
    assert (ss.grad - ss2.grad).abs().max().item() < GTOL
    assert (st.grad - st2.grad).abs().max().item() < GTOL
    for m1 in circ.circuit.layers:
        if isinstance(m1, qandle.operators.BuiltParametrizedOperator) and not m1.named:
            assert m1.theta.grad is not None


def test_c3_family_reduced_batch(device):
    """The adjoint configuration's circuit family (16 qubits, 20 layers, checkpoint every 5 layers) at B = 4."""
    w = workloads.rxryrz_ladder(n=16, B=4, depth=20, seed=2, loss=True)
    gx, gth, _ = grads_fused(w, device)
    ox, oth = oracle_grads(w)
    ex, eth = np.abs(gx - ox).max(), np.abs(gth - oth).max()
    print(f"C3 family B=4: max|dx err|={ex:.2e} max|dtheta err|={eth:.2e} (|dtheta| max {np.abs(oth).max():.2f})")
    assert ex < GTOL and eth < GTOL
