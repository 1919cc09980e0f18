"""Pins of the oracle's remaining operators (oracle/fp64.py apply_gate).

Each check is a fact fixed independently of the restatement: the reference's
own tests of these operators (CZ and SWAP against PennyLane's textbook gates,
src/qandle/test/test_operators.py:109-124, 180-203; Reset's marginals and norm,
:6-55; U with the identity, :71-82) restated as textbook identities, and
closed forms for Controlled and Invert, for which the reference has no test.
"""
import numpy as np
import pytest
import torch

from oracle import fp64

H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2)
X = np.array([[0, 1], [1, 0]], dtype=np.complex128)


def _state(n, B=3, seed=0):
    rng = np.random.default_rng(seed)
    psi = rng.standard_normal((B, 2 ** n)) + 1j * rng.standard_normal((B, 2 ** n))
    return psi / np.linalg.norm(psi, axis=1, keepdims=True)


@pytest.mark.parametrize("n", [2, 3, 4])
def test_cz_is_h_cnot_h(n):
    psi = _state(n)
    for c in range(n):
        for t in range(n):
            if c == t:
                continue
            ref = fp64.apply_1q(fp64.apply_cnot(fp64.apply_1q(psi, n, t, H), n, c, t), n, t, H)
            np.testing.assert_allclose(fp64.apply_gate(psi, n, ("cz", c, t)), ref, atol=1e-12)
            np.testing.assert_array_equal(fp64.apply_gate(psi, n, ("cz", c, t)), fp64.apply_gate(psi, n, ("cz", t, c)))


@pytest.mark.parametrize("n", [2, 3, 4])
def test_swap_is_three_cnots(n):
    psi = _state(n, seed=1)
    for a in range(n):
        for b in range(n):
            out = fp64.apply_gate(psi, n, ("swap", a, b))
            if a == b:
                np.testing.assert_array_equal(out, psi)
                continue
            ref = fp64.apply_cnot(fp64.apply_cnot(fp64.apply_cnot(psi, n, a, b), n, b, a), n, a, b)
            np.testing.assert_array_equal(out, ref)


def test_custom_identity_and_convention():
    n = 3
    psi = _state(n, seed=2)
    np.testing.assert_array_equal(fp64.apply_gate(psi, n, ("u", 1, np.eye(2))), psi)
    # the reference's own RY matrix block M (state @ M) is RY^T; U(M) must act as RY
    phi = 0.7
    M = np.array([[np.cos(phi / 2), np.sin(phi / 2)], [-np.sin(phi / 2), np.cos(phi / 2)]])
    ref = fp64.apply_1q(psi, n, 2, np.array([[np.cos(phi / 2), -np.sin(phi / 2)], [np.sin(phi / 2), np.cos(phi / 2)]]))
    np.testing.assert_allclose(fp64.apply_gate(psi, n, ("u", 2, M)), ref, atol=1e-14)


def test_controlled_x_is_cnot_and_crz_closed_form():
    n = 3
    psi = _state(n, seed=3)
    for c, t in [(0, 1), (2, 0), (1, 2)]:
        np.testing.assert_array_equal(fp64.apply_gate(psi, n, ("ctrl", c, ("u", t, X))), fp64.apply_cnot(psi, n, c, t))
    # CRZ on basis states: phases exp(-+ i phi / 2) only where the control is 1
    phi = 1.3
    for i in range(2 ** n):
        e = np.zeros((1, 2 ** n), dtype=np.complex128)
        e[0, i] = 1
        out = fp64.apply_gate(e, n, ("ctrl", 0, ("rz", 2, phi)))
        cbit, tbit = (i >> 2) & 1, i & 1
        expect = 1.0 if not cbit else np.exp((-0.5j if not tbit else 0.5j) * phi)
        np.testing.assert_allclose(out[0, i], expect, atol=1e-14)
        assert np.abs(out).sum() == pytest.approx(1.0)
    # a nested Controlled contributes only its own target (to_matrix of a Controlled)
    np.testing.assert_array_equal(fp64.apply_gate(psi, n, ("ctrl", 0, ("ctrl", 1, ("u", 2, X)))),
                                  fp64.apply_cnot(psi, n, 0, 2))


def test_invert_undoes():
    n = 3
    psi = _state(n, seed=4)
    rng = np.random.default_rng(5)
    M = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))  # not unitary: a true inverse
    for op in [("rx", 0, 0.4), ("ry", 1, -1.1), ("rz", 2, 2.5), ("u", 1, M), ("cnot", 0, 2), ("cz", 1, 2),
               ("swap", 0, 1), ("ctrl", 0, ("ry", 1, 0.8))]:
        there = fp64.apply_gate(psi, n, op) if op[0] != "ctrl" else fp64.apply_gate(psi, n, op[2])
        back = fp64.apply_gate(there, n, ("inv", op))
        np.testing.assert_allclose(back, psi, atol=1e-12)
    np.testing.assert_allclose(fp64.apply_gate(psi, n, ("inv", ("rx", 0, 0.4))),
                               fp64.apply_1q(psi, n, 0, fp64.rot_matrix("x", -0.4)), atol=1e-14)
    with pytest.raises(ValueError):
        fp64.apply_gate(psi, n, ("inv", ("reset", 0)))


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_reset_marginals_and_norm(n):
    """src/qandle/test/test_operators.py:6-55 restated."""
    idx = np.arange(2 ** n)
    inp = (idx + 0.5 - 0.1j)[None, :].astype(np.complex128)
    inp /= np.linalg.norm(inp)
    batch = _state(n, B=7, seed=n)
    for psi in (inp, batch):
        before = fp64.measure_prob(psi, n)
        for q in range(n):
            out = fp64.apply_gate(psi, n, ("reset", q))
            after = fp64.measure_prob(out, n)
            np.testing.assert_allclose(after[:, q], 1.0, atol=1e-5)  # torch.allclose bar of the reference test
            others = [k for k in range(n) if k != q]
            np.testing.assert_allclose(after[:, others], before[:, others], atol=1e-5)
            np.testing.assert_allclose(np.linalg.norm(out, axis=1), 1.0, atol=1e-5)


def test_reset_keeps_a_reset_qubit():
    n = 3
    psi = _state(n, seed=9)
    zero = fp64.apply_gate(psi, n, ("reset", 1))
    again = fp64.apply_gate(zero, n, ("reset", 1))
    np.testing.assert_allclose(again, zero, rtol=0, atol=1e-6)


def test_lowering_of_swap_and_inverses():
    """SWAP lowers to three CNOTs, CZ to fixed-angle rotations around a CNOT,
    Invert(CNOT/SWAP) to themselves and Invert(rotation) to the rotation with
    a negated slot; the rest stays operator by operator."""
    import warnings

    import qandle_amd as qandle
    from qandle_amd._engine import fused
    from qandle_amd._engine import planner as PL

    n = 4
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        circ = qandle.Circuit([qandle.SWAP(0, 2), qandle.Invert(qandle.CNOT(1, 3)),
                               qandle.Invert(qandle.RY(1, theta=0.5, remapping=None)), qandle.SWAP(1, 1)],
                              num_qubits=n)
    prog = fused.lower_circuit(circ.circuit)
    assert [(g.kind, g.pos, g.ctrl) for g in prog.gates] == [
        (PL.CNOT, 1, 3), (PL.CNOT, 3, 1), (PL.CNOT, 1, 3), (PL.CNOT, 0, 2), (PL.RY, 2, -1)]
    assert prog.slots[prog.gates[-1].slot].sign == -1.0
    cz = qandle.Circuit([qandle.CZ(3, 1)], num_qubits=n)
    p2 = fused.lower_circuit(cz.circuit)
    # CZ(3, 1): target qubit 1 = position 2, control qubit 3 = position 0
    assert [(g.kind, g.pos, g.ctrl) for g in p2.gates] == [
        (PL.RZ, 2, -1), (PL.RY, 2, -1), (PL.CNOT, 2, 0), (PL.RZ, 2, -1), (PL.RY, 2, -1), (PL.RZ, 2, -1)]
    # the decomposition is CZ itself (oracle, fp64)
    ops = [(("rz", "ry", "cnot", "rz", "ry", "rz")[k], 1 if k != 2 else 3, v)
           for k, v in enumerate((np.pi, np.pi / 2, None, np.pi, np.pi / 2, 2 * np.pi))]
    ops = [("cnot", 3, 1) if o[0] == "cnot" else o for o in ops]
    psi = fp64.simulate(n, [("ry", q, 0.3 + q) for q in range(n)])
    np.testing.assert_allclose(fp64.simulate(n, ops, psi), fp64.apply_gate(psi, n, ("cz", 3, 1)), atol=1e-12)
    for extra in (qandle.U(0, torch.eye(2)), qandle.Reset(0), qandle.Controlled(0, qandle.RX(1))):
        c2 = qandle.Circuit([qandle.RX(0), extra], num_qubits=n)
        assert fused.lower_circuit(c2.circuit) is None


def test_modules_move_between_devices_and_keep_parameters():
    """Built operators are torch modules: .to() must keep working (the engine's
    per-gate method must not shadow torch.nn.Module internals) and parameters
    of Invert/Controlled targets appear under the reference's names."""
    import warnings

    import qandle_amd as qandle

    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        layers = [qandle.U(0, torch.eye(2)), qandle.CZ(0, 1), qandle.SWAP(1, 2), qandle.Reset(2),
                  qandle.Invert(qandle.RX(1, theta=0.2)), qandle.Controlled(0, qandle.RY(2, theta=0.4))]
        circ = qandle.Circuit(layers, num_qubits=3).to("cpu").to(torch.float32)
    keys = set(circ.state_dict().keys())
    assert "circuit.layers.4.target.theta" in keys and "circuit.layers.5.t.theta" in keys
    assert circ.circuit.layers[0]._u.dtype == torch.float32
    for m in circ.circuit.layers:
        assert isinstance(str(m), str)


def test_amplitude_embedding_and_joint_probability_pins():
    """AmplitudeEmbedding known answers (3-4-5 row, padding, cropping, unit
    norm) and MeasureJointProbability summing to the squared norm."""
    out = fp64.amplitude_embed(np.array([[3.0, 4.0]]), 2, pad=0, normalize=True)
    np.testing.assert_allclose(out, [[0.6, 0.8, 0.0, 0.0]], atol=1e-15)
    out = fp64.amplitude_embed(np.array([[1.0, 2.0, 3.0]]), 2, pad=7.0, normalize=False)
    np.testing.assert_array_equal(out.real, [[1.0, 2.0, 3.0, 7.0]])
    out = fp64.amplitude_embed(np.arange(10.0)[None, :], 3, pad=0, normalize=False)
    np.testing.assert_array_equal(out.real, [np.arange(8.0)])
    rng = np.random.default_rng(0)
    X = rng.standard_normal((5, 13))
    v = fp64.amplitude_embed(X, 4, pad=0.5, normalize=True)
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-14)
    np.testing.assert_allclose(fp64.amplitude_embed(np.zeros((1, 4)), 2, normalize=True), 0.0)
    with pytest.raises(ValueError):
        fp64.amplitude_embed(X, 4)
    psi = _state(4, seed=3) * 1.7
    np.testing.assert_allclose(fp64.joint_prob(psi).sum(axis=1), 1.7 ** 2, atol=1e-12)
