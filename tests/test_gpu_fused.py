"""GPU parity of the fused program path (Circuit.forward) against the oracle.

Bar: probabilities and amplitudes within 1e-5 absolute of qandle's complex64
semantics (the dense restatement where it exists, n <= 10) and of the FP64
oracle; the FP64 error is printed per configuration.
"""
import numpy as np
import pytest
import torch

import qandle_amd as qandle
from oracle import fp64, qandle_c64
from qandle_amd import workloads
from qandle_amd._engine import fused

pytestmark = pytest.mark.gpu
ATOL = 1e-5


def run_fused(w, device, **kw):
    circ = qandle.Circuit(w.layers, num_qubits=w.n).to(device)
    with torch.no_grad():
        return circ(x=w.inputs["x"].to(device), **kw), circ


def test_c1_family_vs_dense_reference(device):
    w = workloads.embed_ry_ring()  # 8 qubits, B = 32, 4 layers
    P, circ = run_fused(w, device)
    assert fused.lower_circuit(circ.circuit) is not None
    dense = qandle_c64.DenseCircuit(w.n, w.torch_ops())(w.torch_ops())
    assert P.shape == (32, 8)
    assert (P.cpu() - dense).abs().max().item() < ATOL
    ref = fp64.measure_prob(fp64.simulate(w.n, w.oracle_ops()), w.n)
    print("C1 fused: max |P - P_fp64| =", np.abs(P.cpu().numpy() - ref).max())


@pytest.mark.parametrize("n,B,depth", [(5, 13, 3), (9, 7, 2), (12, 4, 3), (13, 3, 2), (14, 2, 3), (16, 2, 2)])
def test_rxryrz_ladder_vs_fp64(device, n, B, depth):
    w = workloads.rxryrz_ladder(n=n, B=B, depth=depth, seed=n)
    P, _ = run_fused(w, device)
    ref = fp64.measure_prob(fp64.simulate(w.n, w.oracle_ops()), w.n)
    err = np.abs(P.cpu().numpy() - ref).max()
    print(f"ladder n={n} B={B} depth={depth}: max |P - P_fp64| = {err:.2e}")
    assert err < ATOL


def test_c2_family_full_width(device):
    """20 qubits, 10 layers (the C2 circuit family) at B = 2: FP64 oracle on the full state."""
    w = workloads.rxryrz_ladder(n=20, B=2, depth=10, seed=1)
    P, circ = run_fused(w, device)
    ref = fp64.measure_prob(fp64.simulate(w.n, w.oracle_ops()), w.n)
    err = np.abs(P.cpu().numpy() - ref).max()
    print(f"C2 family n=20 B=2: max |P - P_fp64| = {err:.2e}")
    assert err < ATOL


def test_state_output_and_input_state(device):
    """No measurement: the final state itself, from a caller-provided input state."""
    gen = torch.Generator().manual_seed(3)
    n, B = 14, 3
    psi0 = torch.randn(B, 2 ** n, dtype=torch.cfloat, generator=gen)
    psi0 = psi0 / torch.linalg.norm(psi0, dim=-1, keepdim=True)
    layers = []
    ops = []
    for q in range(n):
        th = float(torch.rand((), generator=gen) * 6.28)
        layers.append(qandle.RY(q, theta=th, remapping=None))
        ops.append(("ry", q, th))
    for q in range(n):
        layers.append(qandle.CNOT(q, (q + 3) % n))
        ops.append(("cnot", q, (q + 3) % n))
        th = float(torch.rand((), generator=gen) * 6.28)
        layers.append(qandle.RZ(q, theta=th, remapping=None))
        ops.append(("rz", q, th))
    circ = qandle.Circuit(layers, num_qubits=n).to(device)
    with torch.no_grad():
        out = circ(psi0.to(device))
        out1 = circ(psi0[0].to(device))
    ref = fp64.simulate(n, [(k, q, np.float64(np.float32(a))) if k != "cnot" else (k, q, a) for k, q, a in ops],
                        psi0.numpy().astype(np.complex128))
    assert out.shape == (B, 2 ** n) and out1.shape == (2 ** n,)
    assert np.abs(out.cpu().numpy() - ref).max() < ATOL
    assert np.abs(out1.cpu().numpy() - ref[0]).max() < ATOL
    assert torch.equal(psi0, psi0.clone())  # caller's state untouched (out-of-place)


def test_named_per_sample_inputs(device):
    """Re-uploading named gates with (B,) inputs inside the fused program."""
    gen = torch.Generator().manual_seed(5)
    n, B = 10, 6
    x = torch.rand(B, n, generator=gen) * 3
    r = torch.rand(B, generator=gen) * 3
    layers = [qandle.AngleEmbedding(name="x", rotation="ry"), qandle.RX(2, name="r", remapping=None),
              qandle.CNOT(2, 7), qandle.RZ(7, name="r"), qandle.MeasureProbability()]
    circ = qandle.Circuit(layers, num_qubits=n).to(device)
    with torch.no_grad():
        P = circ(x=x.to(device), r=r.to(device))
        P0 = circ(x=x.to(device), r=r[0].to(device))  # 0-d named input: shared
    rr = r.double().numpy()
    ops = [("embed", "y", list(range(n)), x.double().numpy()), ("rx", 2, rr), ("cnot", 2, 7), ("rz", 7, rr)]
    ref = fp64.measure_prob(fp64.simulate(n, ops), n)
    assert np.abs(P.cpu().numpy() - ref).max() < ATOL
    ops0 = [("embed", "y", list(range(n)), x.double().numpy()), ("rx", 2, rr[0]), ("cnot", 2, 7), ("rz", 7, rr[0])]
    assert np.abs(P0.cpu().numpy() - fp64.measure_prob(fp64.simulate(n, ops0), n)).max() < ATOL


def test_split_circuit_equals_unsplit(device):
    """``split_max_qubits`` / ``Circuit.split`` give the unsplit result (``test_splitter.py:55-101``)."""
    torch.manual_seed(0)
    n = 10
    layers = []
    for q in range(n):
        layers.append(qandle.RY(q, remapping=None))
    for q in range(n - 1):
        layers.append(qandle.CNOT(q, q + 1))
    for q in range(n):
        layers.append(qandle.RX(q, remapping=None))
    c = qandle.Circuit(layers, num_qubits=n).to(device)
    s = c.split(max_qubits=4).to(device)
    assert isinstance(s.circuit, qandle.SplittedCircuit)
    gen = torch.Generator().manual_seed(1)
    inp = torch.rand(2 ** n, dtype=torch.cfloat, generator=gen)
    inp = (inp / torch.linalg.norm(inp)).to(device)
    with torch.no_grad():
        a, b = c(inp), s(inp)
        b_layers = s.circuit.forward_layers(inp)
    assert torch.allclose(a, b, atol=1e-6)
    assert torch.allclose(a, b_layers, atol=1e-6)


@pytest.mark.parametrize("max_qubits", [2, 3, 4])
def test_split_decomposed_su_equals_unsplit(device, max_qubits):
    """SU's decomposition split into groups of 2, 3 and 4 qubits gives the
    unsplit state (``test_splitter.py:84-101``), and every gate lands in one
    sub-circuit (``:68-76``)."""
    torch.manual_seed(4)
    su = qandle.SU(qubits=list(range(5)), reps=2, rotations=["rx", "ry"])
    c = qandle.Circuit(layers=su.decompose()).decompose()
    s = c.split(max_qubits=max_qubits)
    assert len(c.circuit.layers) == sum(len(sc.layers) for sc in s.circuit.subcircuits)
    c, s = c.to(device), s.to(device)
    gen = torch.Generator().manual_seed(5)
    inp = torch.rand(2 ** 5, dtype=torch.cfloat, generator=gen)
    inp = (inp / torch.linalg.norm(inp)).to(device)
    with torch.no_grad():
        a, b = c(inp), s(inp)
    ops = []
    for m in c.circuit.layers:
        if isinstance(m, qandle.operators.BuiltCNOT):
            ops.append(("cnot", m.c, m.t))
        else:
            kind = {qandle.operators.BuiltRX: "rx", qandle.operators.BuiltRY: "ry"}[type(m)]
            ops.append((kind, m.qubit, float(m.remapping(m.theta).detach().reshape(-1)[0])))
    ref = fp64.simulate(5, ops, inp.cpu().numpy().astype(np.complex128)[None, :])[0]
    assert np.abs(a.cpu().numpy() - ref).max() < ATOL
    assert torch.allclose(a, b, atol=1e-6)


def test_decomposed_measurement_returns_the_state(device):
    """A decomposed circuit's measurement is a no-op (``measurements.py:24-54``)."""
    import warnings

    torch.manual_seed(6)
    n = 7
    body = [qandle.RY(q, remapping=None) for q in range(n)] + [qandle.CNOT(q, (q + 1) % n) for q in range(n)]
    circ = qandle.Circuit(body + [qandle.MeasureProbability()], num_qubits=n)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dec = circ.decompose().to(device)
        gen = torch.Generator().manual_seed(7)
        inp = torch.randn(3, 2 ** n, dtype=torch.cfloat, generator=gen)
        inp = (inp / torch.linalg.norm(inp, dim=-1, keepdim=True)).to(device)
        with torch.no_grad():
            out = dec(inp)
            plain = dec.circuit.forward_layers(inp)
    assert out.shape == (3, 2 ** n) and out.dtype == torch.complex64
    assert (out - plain).abs().max().item() < ATOL
    ops = [("ry", m.qubit, float(m.theta.detach().reshape(-1)[0])) for m in circ.circuit.layers[:n]]
    ops += [("cnot", q, (q + 1) % n) for q in range(n)]
    ref = fp64.simulate(n, ops, inp.cpu().numpy().astype(np.complex128))
    assert np.abs(out.cpu().numpy() - ref).max() < ATOL


@pytest.mark.parametrize("R,TB", [(5, 7), (5, 6), (4, 8), (3, 8)])
def test_kernel_variants(device, R, TB):
    w = workloads.rxryrz_ladder(n=15, B=3, depth=2, seed=7)
    circ = qandle.Circuit(w.layers, num_qubits=w.n).to(device)
    prog = fused.lower_circuit(circ.circuit)
    with torch.no_grad():
        inp = fused.gather_inputs(prog, circ.circuit, None, {"x": w.inputs["x"].to(device)})
        cs = fused.build_coeffs(prog, inp)
        efac = fused.product_factors(prog, inp)
        psi, P = fused.execute(prog, inp, cs, efac, R, TB)
    ref = fp64.simulate(w.n, w.oracle_ops())
    assert np.abs(psi.cpu().numpy() - ref).max() < ATOL
    assert np.abs(P.cpu().numpy() - fp64.measure_prob(ref, w.n)).max() < ATOL


@pytest.mark.parametrize("n,B,seed", [(15, 3, 1), (17, 2, 2)])
def test_random_long_range_cnots_vs_fp64(device, n, B, seed):
    """Random rotations and long-range CNOTs (any control, any target) through the
    compiled kernels and the lookahead planner, against the FP64 oracle."""
    rng = np.random.default_rng(seed)
    gen = torch.Generator().manual_seed(seed)
    x = torch.rand(B, n, generator=gen) * 3
    layers = [qandle.AngleEmbedding(name="x", qubits=list(range(n)))]
    ops = [("embed", "x", list(range(n)), x.double().numpy())]
    for _ in range(6):
        for q in range(n):
            k = ("rx", "ry", "rz")[int(rng.integers(3))]
            th = float(rng.uniform(0, 2 * np.pi))
            layers.append({"rx": qandle.RX, "ry": qandle.RY, "rz": qandle.RZ}[k](q, theta=th, remapping=None))
            ops.append((k, q, th))
        for _ in range(n // 2):
            c, t = (int(v) for v in rng.choice(n, 2, replace=False))
            layers.append(qandle.CNOT(c, t))
            ops.append(("cnot", c, t))
    layers.append(qandle.MeasureProbability())
    circ = qandle.Circuit(layers, num_qubits=n).to(device)
    assert fused.lower_circuit(circ.circuit) is not None
    with torch.no_grad():
        P = circ(x=x.to(device)).cpu().numpy()
    ref = fp64.measure_prob(fp64.simulate(n, ops), n)
    print(f"random n={n}: max|P err| {np.abs(P - ref).max():.2e}")
    assert np.abs(P - ref).max() < ATOL
