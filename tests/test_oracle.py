"""Pins the CPU oracle (``oracle/``) with checks that do not restate its formulas.

* gate closed forms vs scipy.linalg.expm(-i phi/2 P) (the textbook definition
  PennyLane uses in the reference's differential tests,
  ``src/qandle/test/test_operators.py:58-128``)
* reshape-based application vs explicit Kronecker products (wire 0 most significant)
* CNOT vs a permutation built from per-wire bit lists; involution
* AngleEmbedding vs the product of expm columns; P(q=0) = cos^2(x/2)
  (mirrors ``src/qandle/test/test_embeddings.py:56-153``)
* MeasureProbability vs an explicit loop; P0 + P1 = ||psi||^2
  (mirrors ``src/qandle/test/test_measurements.py:6-48``)
* FP64 adjoint vs central finite differences and parameter shift
* qandle_c64 (dense reference semantics) vs fp64, and dense == structured
"""
import numpy as np
import pytest
import scipy.linalg
import torch

from oracle import fp64, qandle_c64

PAULI = fp64.PAULI
V_UNBATCHED = 0.543321  # src/qandle/test/test_operators.py:87
V_BATCHED = -2.3456     # src/qandle/test/test_operators.py:159


def expm_rot(kind, phi):
    return scipy.linalg.expm(-1j * phi / 2 * PAULI[kind])


@pytest.mark.parametrize("kind", "xyz")
@pytest.mark.parametrize("phi", [V_UNBATCHED, V_BATCHED, 0.0, np.pi, 3.9])
def test_rot_matrix_is_expm(kind, phi):
    np.testing.assert_allclose(fp64.rot_matrix(kind, phi), expm_rot(kind, phi), atol=1e-14)


def test_rot_known_answers():
    phi = 0.7
    c, s = np.cos(phi / 2), np.sin(phi / 2)
    zero = np.array([1, 0], dtype=complex)
    np.testing.assert_allclose(fp64.rot_matrix("x", phi) @ zero, [c, -1j * s])
    np.testing.assert_allclose(fp64.rot_matrix("y", phi) @ zero, [c, s])
    np.testing.assert_allclose(np.diag(fp64.rot_matrix("z", phi)), [np.exp(-0.5j * phi), np.exp(0.5j * phi)])
    for k in "xyz":
        np.testing.assert_allclose(fp64.rot_matrix(k, phi) @ fp64.rot_matrix(k, -phi), np.eye(2), atol=1e-15)


def _rand_state(rng, B, n):
    v = rng.normal(size=(B, 2 ** n)) + 1j * rng.normal(size=(B, 2 ** n))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


@pytest.mark.parametrize("n", range(1, 7))
def test_apply_1q_matches_kron(n):
    rng = np.random.default_rng(n)
    psi = _rand_state(rng, 3, n)
    for q in range(n):
        for kind in "xyz":
            U = expm_rot(kind, V_BATCHED)
            dense = fp64.kron_1q(n, q, U)
            np.testing.assert_allclose(fp64.apply_1q(psi, n, q, U), psi @ dense.T, atol=1e-13)
            # per-sample matrices
            phis = rng.uniform(0, 2 * np.pi, 3)
            Us = np.stack([expm_rot(kind, p) for p in phis])
            got = fp64.apply_1q(psi, n, q, Us)
            for b in range(3):
                np.testing.assert_allclose(got[b], fp64.kron_1q(n, q, Us[b]) @ psi[b], atol=1e-13)


@pytest.mark.parametrize("n", range(2, 7))
def test_cnot_permutation(n):
    rng = np.random.default_rng(10 + n)
    psi = _rand_state(rng, 2, n)
    for c in range(n):
        for t in range(n):
            if c == t:
                continue
            M = fp64.cnot_matrix(n, c, t)
            got = fp64.apply_cnot(psi, n, c, t)
            np.testing.assert_array_equal(got, psi @ M.T)
            np.testing.assert_array_equal(fp64.apply_cnot(got, n, c, t), psi)


def test_cnot_textbook_two_qubits():
    # CNOT(0 -> 1) on |10> gives |11>  (wire 0 is the most significant bit)
    psi = np.zeros((1, 4), dtype=complex)
    psi[0, 0b10] = 1
    out = fp64.apply_cnot(psi, 2, 0, 1)
    assert out[0, 0b11] == 1 and abs(out).sum() == 1


@pytest.mark.parametrize("rot,n", [("x", 3), ("y", 4), ("z", 5)])
def test_embed_is_product_of_columns(rot, n):
    rng = np.random.default_rng(3)
    X = rng.uniform(0, np.pi, size=(4, n))
    got = fp64.embed(X, n, list(range(n)), rot)
    for b in range(4):
        v = np.ones(1, dtype=complex)
        for j in range(n):
            v = np.kron(v, expm_rot(rot, X[b, j])[:, 0])
        np.testing.assert_allclose(got[b], v, atol=1e-14)
    P = fp64.measure_prob(got, n)
    if rot in "xy":
        np.testing.assert_allclose(P, np.cos(X / 2) ** 2, atol=1e-14)
    else:
        np.testing.assert_allclose(P, np.ones_like(P), atol=1e-14)


def test_embed_subset_leaves_others_zero():
    X = np.array([[0.3, 1.1]])
    psi = fp64.embed(X, 4, [3, 1], "y")
    P = fp64.measure_prob(psi, 4)
    np.testing.assert_allclose(P[0], [1, np.cos(1.1 / 2) ** 2, 1, np.cos(0.3 / 2) ** 2], atol=1e-14)


def test_measure_prob_definition():
    rng = np.random.default_rng(5)
    n = 4
    psi = _rand_state(rng, 17, n) * rng.uniform(0.5, 2, size=(17, 1))
    P = fp64.measure_prob(psi, n)
    for b in range(17):
        for q in range(n):
            ref = sum(abs(psi[b, i]) ** 2 for i in range(2 ** n) if not (i >> (n - 1 - q)) & 1)
            assert abs(P[b, q] - ref) < 1e-12
    # P(q=1) computed independently from the complementary mask
    norms = (abs(psi) ** 2).sum(1)
    P1 = np.stack([(abs(psi[:, [i for i in range(16) if (i >> (n - 1 - q)) & 1]]) ** 2).sum(1) for q in range(n)], 1)
    np.testing.assert_allclose(P + P1, np.repeat(norms[:, None], n, 1), atol=1e-12)


def _random_circuit(rng, n, depth, per_sample_B=None):
    ops = [("embed", "x", list(range(n)), rng.uniform(0, np.pi, size=(per_sample_B or 2, n)))]
    for _ in range(depth):
        for q in range(n):
            k = "r" + "xyz"[rng.integers(3)]
            ops.append((k, q, float(rng.uniform(0, 2 * np.pi))))
        for q in range(n - 1):
            ops.append(("cnot", q, q + 1))
    if per_sample_B:
        ops.append(("ry", 0, rng.uniform(0, 2 * np.pi, size=per_sample_B)))
    return ops


@pytest.mark.parametrize("n", [2, 3, 5])
def test_dense_unitary_matches_simulate(n):
    rng = np.random.default_rng(20 + n)
    ops = [op for op in _random_circuit(rng, n, 3) if op[0] != "embed"]
    psi0 = _rand_state(rng, 2, n)
    U = fp64.dense_unitary(n, ops)
    np.testing.assert_allclose(fp64.simulate(n, ops, psi0), psi0 @ U.T, atol=1e-12)
    np.testing.assert_allclose(U.conj().T @ U, np.eye(2 ** n), atol=1e-12)


@pytest.mark.parametrize("n", [3, 4])
def test_adjoint_matches_fd_and_shift(n):
    rng = np.random.default_rng(30 + n)
    B = 3
    ops = _random_circuit(rng, n, 2, per_sample_B=B)
    w = rng.normal(size=(B, n))
    adj = fp64.adjoint_grads(n, ops, w)
    for k, op in enumerate(ops):
        if op[0] == "cnot":
            assert adj[k] is None
            continue
        fd = fp64.finite_diff_grad(n, ops, w, k)
        np.testing.assert_allclose(adj[k], fd, atol=1e-7)
        if op[0] != "embed" or True:
            ps = fp64.param_shift_grad(n, ops, w, k)
            np.testing.assert_allclose(adj[k], ps, atol=1e-12)


def test_planted_mistake_is_caught(monkeypatch):
    """A transposed RY (the classic row/column convention slip) must fail the expm pin."""
    good = fp64.rot_matrix

    def transposed(kind, phi):
        return np.swapaxes(good(kind, phi), -1, -2)

    monkeypatch.setattr(fp64, "rot_matrix", transposed)
    assert not np.allclose(fp64.rot_matrix("y", 0.7), expm_rot("y", 0.7))
    # ... and the gradient check catches a dropped factor 1/2 in the adjoint
    monkeypatch.setattr(fp64, "rot_matrix", good)
    rng = np.random.default_rng(1)
    ops = _random_circuit(rng, 3, 1)
    w = rng.normal(size=(2, 3))
    adj = fp64.adjoint_grads(3, ops, w)
    fd = fp64.finite_diff_grad(3, ops, w, 1)
    assert not np.allclose(2 * adj[1], fd, atol=1e-6)


# ------------------------------------------------------------ qandle_c64

def _to_torch_ops(ops):
    out = []
    for op in ops:
        if op[0] == "embed":
            out.append((op[0], op[1], op[2], torch.tensor(op[3], dtype=torch.float32)))
        elif op[0] == "cnot":
            out.append(op)
        else:
            out.append((op[0], op[1], torch.tensor(op[2], dtype=torch.float32)))
    return out


def _f32_ops(ops):
    """fp64 ops with angles rounded to float32, so both sides see the same inputs."""
    out = []
    for op in ops:
        if op[0] == "embed":
            out.append((op[0], op[1], op[2], np.asarray(op[3], np.float32).astype(np.float64)))
        elif op[0] == "cnot":
            out.append(op)
        else:
            out.append((op[0], op[1], np.float32(op[2]).astype(np.float64) if np.ndim(op[2]) == 0 else np.asarray(op[2], np.float32).astype(np.float64)))
    return out


@pytest.mark.parametrize("n", [1, 3, 5])
def test_c64_dense_matches_fp64(n):
    rng = np.random.default_rng(40 + n)
    B = 4
    ops = _random_circuit(rng, n, 2, per_sample_B=B)
    tops = _to_torch_ops(ops)
    dense = qandle_c64.DenseCircuit(n, tops)
    P = dense(tops).numpy()
    ref = fp64.measure_prob(fp64.simulate(n, _f32_ops(ops)), n)
    np.testing.assert_allclose(P.reshape(ref.shape), ref, atol=2e-6)
    st = dense(tops, measure=False)
    st_s = qandle_c64.structured_run(n, tops, B, measure=False)
    np.testing.assert_allclose(st.numpy(), st_s.numpy(), atol=2e-6)


def test_c64_measure_quirks():
    n = 3
    s = torch.randn(1, 8, dtype=torch.complex64)
    assert qandle_c64.dense_measure_prob(s, n).shape == (n,)
    assert qandle_c64.dense_measure_prob(s[0], n).shape == (n,)
    assert qandle_c64.dense_measure_prob(torch.randn(5, 8, dtype=torch.complex64), n).shape == (5, n)


def test_c64_reference_known_answers():
    """RX/RY/RZ(v) on random states vs expm, the reference's own test values."""
    torch.manual_seed(42)
    for n in range(1, 6):
        inp = torch.rand(2 ** n, dtype=torch.cfloat)
        inp = inp / torch.linalg.norm(inp)
        for q in range(n):
            for kind in "xyz":
                g = qandle_c64.DenseGate(kind, n, q)
                got = g(inp, torch.tensor(V_UNBATCHED)).numpy()
                ref = fp64.kron_1q(n, q, expm_rot(kind, np.float32(V_UNBATCHED))) @ inp.numpy().astype(complex)
                assert np.abs(got - ref).sum() < 1e-5
