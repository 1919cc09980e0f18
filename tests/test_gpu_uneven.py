"""Uneven batch (many circuits of different widths and depths in one captured
graph) against the single-circuit fused path and the FP64 oracle."""
import numpy as np
import pytest
import torch

import qandle_amd as qandle
from oracle import fp64
from qandle_amd import workloads
from qandle_amd._engine.uneven import UnevenBatch

pytestmark = pytest.mark.gpu


def make(num, n_lo, n_hi, d_lo, d_hi, seed, device):
    ws = workloads.random_uneven(num, n_lo, n_hi, d_lo, d_hi, 0.3, seed)
    circs = [qandle.Circuit(w.layers, num_qubits=w.n).to(device) for w in ws]
    return ws, circs


@pytest.mark.parametrize("R,TB", [(4, 7), (4, 8), (5, 6), (5, 8)])
def test_uneven_forward_backward_vs_single(device, R, TB):
    ws, circs = make(40, max(4, R), 14, 2, 8, 21, device)
    ub = UnevenBatch(circs, [w.inputs["x"] for w in ws], device, R=R, TB=TB, grid=512)
    gen = torch.Generator().manual_seed(5)
    lw = [torch.randn(w.B, w.n, generator=gen) for w in ws]
    ub.set_loss_weights(lw)
    ub.run_forward()
    ub.run_backward()
    torch.cuda.synchronize()
    for i, (w, c) in enumerate(zip(ws, circs)):
        x = w.inputs["x"].to(device).requires_grad_(True)
        P = c(x=x)
        (P * lw[i].to(device)).sum().backward()
        assert (ub.probs(i).reshape(P.shape) - P.detach()).abs().max().item() < 1e-5
        assert (ub.input_grads(i) - x.grad).abs().max().item() < 1e-4
        th = torch.stack([m.theta.grad.reshape(()) for m in c.circuit.layers
                          if isinstance(m, qandle.operators.BuiltParametrizedOperator)])
        assert (ub.param_grads(i) - th).abs().max().item() < 1e-4


def test_uneven_graph_replay_vs_fp64(device):
    ws, circs = make(24, 4, 16, 2, 12, 33, device)
    ub = UnevenBatch(circs, [w.inputs["x"] for w in ws], device)
    gen = torch.Generator().manual_seed(6)
    lw = [torch.randn(w.B, w.n, generator=gen) for w in ws]
    ub.set_loss_weights(lw)
    ub.capture(backward=True)
    for rep in range(3):  # replays with eager work between them
        ub.replay(backward=True)
        torch.cuda.synchronize()
        _ = torch.ones(7, device=device).sum().item()
        for i in (0, 5, 11, 23):
            w = ws[i]
            ops = w.oracle_ops()
            ref = fp64.measure_prob(fp64.simulate(w.n, ops), w.n)
            assert np.abs(ub.probs(i).cpu().numpy() - ref).max() < 1e-5
            adj = fp64.adjoint_grads(w.n, ops, lw[i].double().numpy())
            assert np.abs(ub.input_grads(i).cpu().numpy() - adj[0]).max() < 1e-4
            gates = np.array([a for a, op in zip(adj[1:], ops[1:]) if op[0] != "cnot"])
            assert np.abs(ub.param_grads(i).cpu().numpy() - gates).max() < 1e-4
