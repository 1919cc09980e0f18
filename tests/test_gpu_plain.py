"""GPU parity of the plain per-gate path (one kernel per operator) against the oracle.

Mirrors the reference's operator tests (``src/qandle/test/test_operators.py:85-203``,
``test_measurements.py:6-48``, ``test_embeddings.py:56-153``) with the FP64
oracle in place of PennyLane.  Tolerances: amplitudes/probabilities 1e-5
absolute (the bar for qandle's complex64 semantics); CNOT bit-exact.
"""
import numpy as np
import pytest
import torch

import qandle_amd as qandle
from oracle import fp64, qandle_c64

pytestmark = pytest.mark.gpu
ATOL = 1e-5


def rand_state(gen, B, n, device):
    v = torch.rand(B, 2 ** n, dtype=torch.cfloat, generator=gen)
    v = v / torch.linalg.norm(v, dim=-1, keepdim=True)
    return v.to(device)


def oracle_rot(psi, n, q, kind, phi):
    """fp64 oracle on float32-rounded angles."""
    phi = np.asarray(phi, np.float32).astype(np.float64)
    return fp64.apply_1q(psi.cpu().numpy().astype(np.complex128), n, q, fp64.rot_matrix(kind, phi))


@pytest.mark.parametrize("v", [0.543321, -2.3456])
def test_rotations_all_qubits(device, v):
    gen = torch.Generator().manual_seed(42)
    for n in range(1, 7):
        psi = rand_state(gen, 13, n, device)
        for q in range(n):
            for cls, kind in ((qandle.RX, "x"), (qandle.RY, "y"), (qandle.RZ, "z")):
                g = cls(qubit=q, theta=v, remapping=None).build(num_qubits=n).to(device)
                out = g(psi)
                ref = oracle_rot(psi, n, q, kind, v)
                assert np.abs(out.detach().cpu().numpy() - ref).max() < ATOL
                out1 = g(psi[0])  # unbatched
                assert out1.shape == (2 ** n,)
                assert np.abs(out1.detach().cpu().numpy() - ref[0]).max() < ATOL


def test_rotation_large_bits(device):
    gen = torch.Generator().manual_seed(1)
    n = 16
    psi = rand_state(gen, 3, n, device)
    for q in (0, 1, 7, 14, 15):
        g = qandle.RY(qubit=q, theta=1.234, remapping=None).build(n).to(device)
        ref = oracle_rot(psi, n, q, "y", 1.234)
        assert np.abs(g(psi).detach().cpu().numpy() - ref).max() < ATOL


def test_cnot_bit_exact(device):
    gen = torch.Generator().manual_seed(2)
    for n in range(2, 7):
        psi = rand_state(gen, 5, n, device)
        for c in range(n):
            for t in range(n):
                if c == t:
                    continue
                g = qandle.CNOT(c, t).build(n)
                out = g(psi)
                ref = qandle_c64.structured_cnot(psi.cpu(), n, c, t)
                assert torch.equal(out.cpu(), ref)
                assert torch.equal(g(out), psi)


def test_reuploading(device):
    """named input == fixed theta, batched and unbatched (``test_operators.py:131-154``)."""
    q = 3
    v = torch.tensor(0.123)
    c1 = qandle.Circuit(layers=[qandle.RX(qubit=1, theta=v, remapping=None)], num_qubits=q).to(device)
    c2 = qandle.Circuit(layers=[qandle.RX(qubit=1, name="reupload", remapping=None)], num_qubits=q).to(device)
    gen = torch.Generator().manual_seed(42)
    inp = rand_state(gen, 1, q, device)[0]
    assert torch.allclose(c1(inp), c2(inp, reupload=v.to(device)), atol=1e-6)
    inpb = rand_state(gen, 7, q, device)
    b1 = c1(inpb)
    assert torch.allclose(b1, c2(inpb, reupload=v.to(device)), atol=1e-6)
    vb = torch.full((7,), 0.123, device=device)
    assert torch.allclose(b1, c2(inpb, reupload=vb), atol=1e-6)
    out = c2(inp, reupload=vb)  # unbatched state + batched input -> batched output
    assert out.shape == (7, 2 ** q)


@pytest.mark.parametrize("rot,n", [("x", 3), ("y", 4), ("z", 5)])
@pytest.mark.parametrize("batched", [False, True])
def test_angle_embedding(device, rot, n, batched):
    gen = torch.Generator().manual_seed(7)
    x = (torch.rand(10, n, generator=gen) + 0.5) if batched else torch.arange(n, dtype=torch.float32) + 2.5
    emb = qandle.AngleEmbedding(qubits=list(range(n)), rotation="r" + rot, name="e").build(n)
    out = emb(e=x.to(device)).cpu().numpy()
    ref = fp64.embed(x.double().numpy(), n, list(range(n)), rot)
    if not batched:
        ref = ref[0]
    assert out.shape == ref.shape
    assert np.abs(out - ref).max() < ATOL


def test_angle_embedding_subset_and_large(device):
    gen = torch.Generator().manual_seed(8)
    x = torch.rand(3, 4, generator=gen) * 3
    emb = qandle.AngleEmbedding(qubits=[5, 0, 12, 3], rotation="Y", name="e").build(14)
    out = emb(e=x.to(device)).cpu().numpy()
    ref = fp64.embed(x.double().numpy(), 14, [5, 0, 12, 3], "y")
    assert np.abs(out - ref).max() < ATOL


@pytest.mark.parametrize("n,B", [(3, 1), (3, 17), (12, 3), (13, 2), (18, 2)])
def test_measure_probability(device, n, B):
    gen = torch.Generator().manual_seed(n * 100 + B)
    psi = rand_state(gen, B, n, device)
    m = qandle.MeasureProbability().build(n)
    out = m(psi).cpu().numpy()
    ref = fp64.measure_prob(psi.cpu().numpy().astype(np.complex128), n)
    ref = ref[0] if B == 1 else ref  # the reference squeezes B = 1
    assert out.shape == ref.shape
    assert np.abs(out - ref).max() < ATOL
    assert m(psi[0]).shape == (n,)


def test_circuit_layers_c1_family_vs_dense_reference(device):
    """8 qubits, B=32: the reference's own dense complex64 path is the comparator."""
    from qandle_amd import workloads

    w = workloads.embed_ry_ring()
    circ = qandle.Circuit(w.layers, num_qubits=w.n).to(device)
    out = circ.circuit.forward_layers(x=w.inputs["x"].to(device)).cpu()
    dense = qandle_c64.DenseCircuit(w.n, w.torch_ops())(w.torch_ops())
    assert out.shape == dense.shape == (32, 8)
    assert (out - dense).abs().max().item() < ATOL
    ref64 = fp64.measure_prob(fp64.simulate(w.n, w.oracle_ops()), w.n)
    print("C1 plain path: max |P - P_fp64| =", np.abs(out.detach().numpy() - ref64).max())


def test_plain_gradients_vs_fp64_adjoint(device):
    from qandle_amd import workloads

    w = workloads.rxryrz_ladder(n=5, B=3, depth=2, seed=11, loss=True)
    circ = qandle.Circuit(w.layers, num_qubits=w.n).to(device)
    x = w.inputs["x"].to(device).requires_grad_(True)
    P = circ.circuit.forward_layers(x=x)
    loss = (P * w.loss_w.to(device)).sum()
    loss.backward()
    ops = w.oracle_ops()
    adj = fp64.adjoint_grads(w.n, ops, w.loss_w.double().numpy())
    assert np.abs(x.grad.cpu().numpy() - adj[0]).max() < 1e-4
    thetas = [m for m in circ.circuit.layers if isinstance(m, qandle.operators.BuiltParametrizedOperator)]
    gate_grads = [a for a, op in zip(adj[1:], ops[1:]) if op[0] != "cnot"]
    got = np.array([m.theta.grad.item() for m in thetas])
    assert np.abs(got - np.array(gate_grads)).max() < 1e-4


def test_roundtrip_and_norm_at_20_qubits(device):
    """Size-independent properties at full width: U^dag U = 1, CNOT^2 = 1 exactly, norm kept."""
    from qandle_amd._engine import plain

    gen = torch.Generator().manual_seed(5)
    n = 20
    psi = rand_state(gen, 2, n, device)
    phi = torch.tensor(0.77, device=device)
    cs = plain.half_angle_coeffs(phi)
    lib = plain.L.load()
    out = torch.empty_like(psi)
    for bit in (0, 9, 19):
        for kind in (0, 1, 2):
            plain.L.check(lib.qs_rot(out.data_ptr(), psi.data_ptr(), 2, n, bit, kind, cs.data_ptr(), 0, 0, 0), "rot")
            back = torch.empty_like(psi)
            plain.L.check(lib.qs_rot(back.data_ptr(), out.data_ptr(), 2, n, bit, kind, cs.data_ptr(), 0, 1, 0), "rot")
            assert (back - psi).abs().max().item() < 1e-6
            n_in, n_out = torch.linalg.norm(psi[0]).item(), torch.linalg.norm(out[0]).item()
            assert abs(n_out - n_in) < 1e-5 * n_in
    x = psi.clone()
    plain.L.check(lib.qs_cnot(x.data_ptr(), x.data_ptr(), 2, n, 19, 0, 0), "cnot")
    plain.L.check(lib.qs_cnot(x.data_ptr(), x.data_ptr(), 2, n, 19, 0, 0), "cnot")
    torch.cuda.synchronize()
    assert torch.equal(x, psi)
