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


def _layout(sb):
    out = []
    for g in sb.layers_ub:
        if hasattr(g, "qubit"):
            out.append((type(g).__name__, g.qubit))
        else:
            out.append((type(g).__name__, g.c, g.t))
    return out


def test_budget_layout_edge_cases():
    """Walks of the generator at src/qandle/ansaetze/stronglyentangling_budget.py:50-75,
    done by hand: a budget that fills whole layers still opens the next layer
    (its ring is kept, it places no rotation); spacing 2 skips every other ring;
    the ring offset is (k mod Q) mod (Q - 1) for the k-th ring."""
    rz, ry, cx = "RZ", "RY", "CNOT"
    # 3 qubits, budget 6, a ring before every layer
    got = _layout(qandle.StronglyEntanglingLayerBudget(3, param_budget=6))
    assert got == [(rz, 0), (rz, 1), (rz, 2), (cx, 0, 1), (cx, 1, 2), (cx, 2, 0),
                   (ry, 0), (ry, 1), (ry, 2), (cx, 0, 2), (cx, 1, 0), (cx, 2, 1)]
    # 3 qubits, budget 7, a ring before every second layer
    got = _layout(qandle.StronglyEntanglingLayerBudget(3, param_budget=7, control_gate_spacing=2))
    assert got == [(rz, 0), (rz, 1), (rz, 2), (ry, 0), (ry, 1), (ry, 2),
                   (cx, 0, 1), (cx, 1, 2), (cx, 2, 0), (rz, 0)]
    # 4 qubits, six rings: offsets 0, 1, 2, 0, 0, 1
    sb = qandle.StronglyEntanglingLayerBudget(4, param_budget=24, rotations=["ry"])
    rings = [g for g in sb.layers_ub if not hasattr(g, "qubit")]
    offsets = [(g.t - g.c - 1) % 4 for g in rings[0::4]]
    assert offsets == [0, 1, 2, 0, 0, 1]
    assert all((rings[4 * k + i].c, rings[4 * k + i].t) == (i, (i + 1 + offsets[k]) % 4)
               for k in range(6) for i in range(4))
    # no budget, no gates; one qubit cannot form a ring
    assert qandle.StronglyEntanglingLayerBudget(3, param_budget=0).layers_ub == []
    with pytest.raises(ZeroDivisionError):
        qandle.StronglyEntanglingLayerBudget(1, param_budget=2)
    with pytest.raises(AssertionError):
        qandle.StronglyEntanglingLayerBudget(3, param_budget=2, control_gate=qandle.SWAP)


def test_sel_decompositions():
    """Built and unbuilt SEL list the same gates when the layer spans the register;
    on a sub-register the unbuilt form's ring offset runs modulo the register size
    minus one (src/qandle/ansaetze/stronglyentangling.py:75 vs :104)."""
    q = torch.rand(4, 3, 3)
    full = qandle.StronglyEntanglingLayer(qubits=[0, 1, 2], num_qubits_total=3, depth=4, q_params=q)
    def sig(gates):
        return [(type(g).__name__.replace("Built", ""), getattr(g, "qubit", None), getattr(g, "c", None),
                 getattr(g, "t", None)) for g in gates]

    unb = full.decompose()
    assert all(isinstance(g, qandle.operators.BuiltCNOT) for g in unb if hasattr(g, "c"))  # as the reference
    assert sig(unb) == sig(full.build(num_qubits=3).decompose())
    sub = qandle.StronglyEntanglingLayer(qubits=[1, 2, 3], num_qubits_total=5, depth=3, q_params=q[:3])
    built = [g for g in sub.build(num_qubits=5).mods if isinstance(g, qandle.operators.BuiltCNOT)]
    assert [(g.c, g.t) for g in built[6:9]] == [(1, 2), (2, 3), (3, 1)]    # offset 2 % 2 = 0
    # unbuilt: offset 2 % 4 = 2 turns the 3-wire ring onto itself, and CNOT(q, q) is refused
    with pytest.raises(AssertionError):
        sub.decompose()
    sub2 = qandle.StronglyEntanglingLayer(qubits=[1, 2, 3], num_qubits_total=5, depth=2, q_params=q[:2])
    rings = [g for g in sub2.decompose() if isinstance(g, qandle.operators.BuiltCNOT)]
    assert [(g.c, g.t) for g in rings] == [(1, 2), (2, 3), (3, 1), (1, 3), (2, 1), (3, 2)]  # offsets 0, 1
    with pytest.raises(TypeError):
        qandle.StronglyEntanglingLayer(qubits=[0, 1], depth=1).decompose()  # no register size given
