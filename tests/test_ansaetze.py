"""Ansätze (src/qandle/ansaetze/): structure, parameter draws, lowering.

The gate order of StronglyEntanglingLayer is pinned against PennyLane's
published StronglyEntanglingLayers definition (per layer Rot(w0, w1, w2) =
RZ(w2) RY(w1) RZ(w0) on every wire, then CNOT(i, (i + r) mod n) with range
r = l mod (n - 1) + 1), the reference test's ground truth
(src/qandle/test/test_ansaetze.py:6-60), through the fp64 oracle.  GPU parity
is in test_gpu_ansaetze.py.
"""
import warnings

import numpy as np
import pytest
import torch

import qandle_amd as qandle
from oracle import fp64
from qandle_amd._engine import fused
from qandle_amd._engine import planner as PL

KIND = {qandle.operators.BuiltRX: "rx", qandle.operators.BuiltRY: "ry", qandle.operators.BuiltRZ: "rz"}


def _ops_of(mods):
    ops = []
    for m in mods:
        if isinstance(m, qandle.operators.BuiltCNOT):
            ops.append(("cnot", m.c, m.t))
        elif isinstance(m, qandle.operators.BuiltCZ):
            ops.append(("cz", m.c, m.t))
        else:
            ops.append((KIND[type(m)], m.qubit, float(m.remapping(m.theta).reshape(-1)[0])))
    return ops


def _pennylane_sel(weights):
    depth, n, _ = weights.shape
    ops = []
    for lay in range(depth):
        for i in range(n):
            w0, w1, w2 = (float(v) for v in weights[lay, i])
            ops += [("rz", i, w0), ("ry", i, w1), ("rz", i, w2)]
        r = lay % (n - 1) + 1 if n > 1 else 0
        ops += [("cnot", i, (i + r) % n) for i in range(n)]
    return ops


@pytest.mark.parametrize("n,depth", [(4, 10), (5, 7), (2, 3)])
def test_sel_gate_order_is_pennylanes(n, depth):
    gen = torch.Generator().manual_seed(n)
    weights = torch.rand(depth, n, 3, generator=gen)
    sel = qandle.StronglyEntanglingLayer(qubits=list(range(n)), depth=depth, q_params=weights,
                                         remapping=None).build(num_qubits=n)
    rng = np.random.default_rng(n)
    psi = rng.standard_normal((2, 2 ** n)) + 1j * rng.standard_normal((2, 2 ** n))
    ours = fp64.simulate(n, _ops_of(sel.mods), psi)
    ref = fp64.simulate(n, _pennylane_sel(weights.double().numpy()), psi)
    np.testing.assert_allclose(ours, ref, atol=1e-6)  # float32 weights


def test_parameter_draws_follow_the_reference_order():
    n = 3
    torch.manual_seed(7)
    sel = qandle.StronglyEntanglingLayer(qubits=[0, 1, 2], depth=2).build(num_qubits=n)
    torch.manual_seed(7)
    q = torch.rand(2, 3, 3)
    got = torch.stack([m.theta.reshape(()) for m in sel.mods if hasattr(m, "theta")])
    assert torch.equal(got, q.reshape(-1))
    torch.manual_seed(8)
    tl = qandle.TwoLocal(depth=2).build(num_qubits=n)
    torch.manual_seed(8)
    q = torch.rand(2, n)
    assert torch.equal(torch.stack([m.theta.reshape(()) for lay in tl.mods for m in lay]), q.reshape(-1))
    torch.manual_seed(9)
    su = qandle.SU(reps=2, rotations=["ry", "rz"]).build(num_qubits=n)
    torch.manual_seed(9)
    q = torch.rand(3 * n * 2)
    assert torch.equal(torch.stack([m.theta.reshape(()) for lay in su.layers for m in lay]), q)
    torch.manual_seed(10)
    sb = qandle.StronglyEntanglingLayerBudget(n, param_budget=7)
    torch.manual_seed(10)
    q = torch.rand(7)
    assert torch.equal(torch.stack([m.theta.reshape(()) for m in sb.layers if hasattr(m, "theta")]), q)


def test_budget_layout():
    """The allocation of src/qandle/ansaetze/stronglyentangling_budget.py:50-75 for
    3 qubits, budget 7, a control ring every layer (derived by hand)."""
    sb = qandle.StronglyEntanglingLayerBudget(3, param_budget=7, control_gate=qandle.CZ)
    got = [(type(g).__name__, getattr(g, "qubit", None), getattr(g, "c", None), getattr(g, "t", None))
           for g in sb.layers_ub]
    rz, ry, cz = "RZ", "RY", "CZ"
    expect = [(rz, 0, None, None), (rz, 1, None, None), (rz, 2, None, None),
              (cz, None, 0, 1), (cz, None, 1, 2), (cz, None, 2, 0),
              (ry, 0, None, None), (ry, 1, None, None), (ry, 2, None, None),
              (cz, None, 0, 2), (cz, None, 1, 0), (cz, None, 2, 1),
              (rz, 0, None, None)]
    assert got == expect


def test_lowering():
    n = 6
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sel = qandle.StronglyEntanglingLayer(qubits=list(range(n)), depth=3, remapping=None)
        c = qandle.Circuit([qandle.AngleEmbedding(name="x", qubits=list(range(n))), sel, qandle.MeasureProbability()],
                           num_qubits=n)
    prog = fused.lower_circuit(c.circuit)
    assert prog is not None and len(prog.gates) == 3 * (3 * n + n)
    built = c.circuit.layers[1]
    assert [(g.kind == PL.CNOT) for g in prog.gates] == [isinstance(m, qandle.operators.BuiltCNOT) for m in built.mods]
    for gate in (qandle.CNOT, qandle.CZ):
        sb = qandle.Circuit([qandle.StronglyEntanglingLayerBudget(n, param_budget=20, control_gate=gate)], num_qubits=n)
        assert fused.lower_circuit(sb.circuit) is not None
    for other in (qandle.TwoLocal(depth=2), qandle.SU(reps=1)):
        assert fused.lower_circuit(qandle.Circuit([other], num_qubits=n).circuit) is None
