"""Golden vectors (tests/golden/*.npz, written by make_golden.py from the FP64
oracle): the oracle still reproduces them, and on the GPU the engine matches
them within the complex64 bar."""
import os

import numpy as np
import pytest
import torch

from oracle import fp64

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _c1_ops(g):
    n = 8
    ops = [("embed", "x", list(range(n)), g["X"].astype(np.float64))]
    for l in range(4):
        for q in range(n):
            ops.append(("ry", q, float(g["theta"][l, q])))
        for q in range(n):
            ops.append(("cnot", q, (q + 1) % n))
    return ops


def test_oracle_reproduces_golden():
    g = np.load(os.path.join(HERE, "c1_family.npz"))
    ops = _c1_ops(g)
    psi = fp64.simulate(8, ops)
    np.testing.assert_allclose(psi, g["psi"], atol=1e-6)
    np.testing.assert_allclose(fp64.measure_prob(psi, 8), g["P"], atol=1e-12)
    k = np.load(os.path.join(HERE, "kat_rotations.npz"))
    psi = k["v1_n3_in"].astype(np.complex128)
    np.testing.assert_allclose(fp64.apply_1q(psi, 3, 1, fp64.rot_matrix("y", float(np.float32(0.543321)))),
                               k["v1_n3_q1_ry"], atol=1e-6)


@pytest.mark.gpu
def test_engine_matches_golden(device):
    import qandle_amd as qandle

    g = np.load(os.path.join(HERE, "c1_family.npz"))
    n = 8
    layers = [qandle.AngleEmbedding(name="x", qubits=list(range(n)), rotation="rx")]
    for l in range(4):
        for q in range(n):
            layers.append(qandle.RY(q, theta=torch.tensor(float(g["theta"][l, q])), remapping=None))
        for q in range(n):
            layers.append(qandle.CNOT(q, (q + 1) % n))
    layers.append(qandle.MeasureProbability())
    circ = qandle.Circuit(layers, num_qubits=n).to(device)
    x = torch.tensor(g["X"], device=device, requires_grad=True)
    P = circ(x=x)
    (P * torch.tensor(g["w"], dtype=torch.float32, device=device)).sum().backward()
    assert np.abs(P.detach().cpu().numpy() - g["P"]).max() < 1e-5
    assert np.abs(x.grad.cpu().numpy() - g["grad_X"]).max() < 1e-4
    gt = np.array([m.theta.grad.item() for m in circ.circuit.layers
                   if isinstance(m, qandle.operators.BuiltParametrizedOperator)])
    assert np.abs(gt - g["grad_theta"]).max() < 1e-4
    k = np.load(os.path.join(HERE, "kat_rotations.npz"))
    for tag, v in (("v1", 0.543321), ("v2", -2.3456)):
        for nn in range(1, 6):
            psi = torch.tensor(k[f"{tag}_n{nn}_in"], device=device)
            for q in range(nn):
                for kind, cls in (("x", qandle.RX), ("y", qandle.RY), ("z", qandle.RZ)):
                    out = cls(qubit=q, theta=v, remapping=None).build(nn).to(device)(psi)
                    assert np.abs(out.detach().cpu().numpy() - k[f"{tag}_n{nn}_q{q}_r{kind}"]).max() < 1e-5
