"""GPU parity of the remaining operators (qs_gates.hip) against the fp64 oracle.

Every operator module of src/qandle/operators.py:72-147, 416-699 runs on its
own (the plain per-gate path) on seeded random states: permutations and sign
flips (CZ, SWAP, controlled CNOT/CZ/SWAP) must match bit for bit, the others
to 1e-5; gradients (controlled and inverted rotations, U, Reset) against
central finite differences of the fp64 oracle to 1e-4.
"""
import numpy as np
import pytest
import torch

import qandle_amd as qandle
from oracle import fp64

pytestmark = pytest.mark.gpu
ATOL = 1e-5


def _f32(v):
    return float(np.float32(v))


def _state(n, B, seed):
    gen = torch.Generator().manual_seed(seed)
    psi = torch.randn(B, 2 ** n, dtype=torch.complex64, generator=gen)
    return psi / torch.linalg.norm(psi, dim=-1, keepdim=True)


def _matrix(seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))).astype(np.complex64)


M1, M2 = _matrix(1), _matrix(2)
N = 5


def _cases():
    a, b = _f32(0.7), _f32(-1.1)
    rx = lambda q, v: qandle.RX(q, theta=v, remapping=None)  # noqa: E731
    ry = lambda q, v: qandle.RY(q, theta=v, remapping=None)  # noqa: E731
    rz = lambda q, v: qandle.RZ(q, theta=v, remapping=None)  # noqa: E731
    tm = lambda m: torch.from_numpy(m)  # noqa: E731
    return [
        ("cz", lambda: qandle.CZ(0, 3), ("cz", 0, 3), True),
        ("cz_rev", lambda: qandle.CZ(4, 1), ("cz", 4, 1), True),
        ("swap", lambda: qandle.SWAP(1, 4), ("swap", 1, 4), True),
        ("swap_same", lambda: qandle.SWAP(2, 2), ("swap", 2, 2), True),
        ("u", lambda: qandle.U(2, tm(M1)), ("u", 2, M1.astype(np.complex128)), False),
        ("inv_rx", lambda: qandle.Invert(rx(1, a)), ("inv", ("rx", 1, a)), False),
        ("inv_u", lambda: qandle.Invert(qandle.U(0, tm(M2))), ("inv", ("u", 0, M2.astype(np.complex128))), False),
        ("inv_cnot", lambda: qandle.Invert(qandle.CNOT(0, 2)), ("inv", ("cnot", 0, 2)), True),
        ("inv_cz", lambda: qandle.Invert(qandle.CZ(1, 3)), ("inv", ("cz", 1, 3)), True),
        ("inv_swap", lambda: qandle.Invert(qandle.SWAP(0, 1)), ("inv", ("swap", 0, 1)), True),
        ("ctrl_ry", lambda: qandle.Controlled(0, ry(2, b)), ("ctrl", 0, ("ry", 2, b)), False),
        ("ctrl_rz", lambda: qandle.Controlled(3, rz(0, a)), ("ctrl", 3, ("rz", 0, a)), False),
        ("ctrl_u", lambda: qandle.Controlled(4, qandle.U(1, tm(M1))), ("ctrl", 4, ("u", 1, M1.astype(np.complex128))), False),
        ("toffoli", lambda: qandle.Controlled(3, qandle.CNOT(0, 1)), ("ctrl", 3, ("cnot", 0, 1)), True),
        ("ccz", lambda: qandle.Controlled(2, qandle.CZ(0, 4)), ("ctrl", 2, ("cz", 0, 4)), True),
        ("fredkin", lambda: qandle.Controlled(1, qandle.SWAP(0, 3)), ("ctrl", 1, ("swap", 0, 3)), True),
        ("ctrl_ctrl", lambda: qandle.Controlled(0, qandle.Controlled(1, rx(2, a))),
         ("ctrl", 0, ("ctrl", 1, ("rx", 2, a))), False),
        ("ctrl_inv", lambda: qandle.Controlled(0, qandle.Invert(rz(3, b))), ("ctrl", 0, ("inv", ("rz", 3, b))), False),
        ("reset", lambda: qandle.Reset(2), ("reset", 2), False),
        ("reset_top", lambda: qandle.Reset(0), ("reset", 0), False),
    ]


@pytest.mark.parametrize("name,make,op,exact", _cases(), ids=[c[0] for c in _cases()])
def test_gate_vs_oracle(device, name, make, op, exact):
    with _nowarn():
        mod = make().build(num_qubits=N).to(device)
    psi = _state(N, 3, seed=len(name))
    with torch.no_grad():
        out = mod(psi.to(device)).cpu().numpy()
        single = mod(psi[0].to(device)).cpu().numpy()
    ref = fp64.apply_gate(psi.numpy().astype(np.complex128), N, op)
    print(f"{name}: max |out - fp64| = {np.abs(out - ref).max():.2e}")
    if exact:
        np.testing.assert_array_equal(out.astype(np.complex128), ref)
    else:
        assert np.abs(out - ref).max() < ATOL
    assert single.shape == (2 ** N,)
    np.testing.assert_array_equal(single, out[0])


class _nowarn:
    def __enter__(self):
        import warnings

        self._c = warnings.catch_warnings()
        self._c.__enter__()
        warnings.simplefilter("ignore")

    def __exit__(self, *a):
        return self._c.__exit__(*a)


def test_controlled_named_per_sample(device):
    B = 4
    gen = torch.Generator().manual_seed(3)
    ang = torch.rand(B, generator=gen) * 3
    mod = qandle.Controlled(2, qandle.RY(0, name="w", remapping=None)).build(num_qubits=N).to(device)
    assert mod.named
    psi = _state(N, B, seed=4)
    with torch.no_grad():
        out = mod(psi.to(device), w=ang.to(device)).cpu().numpy()
        one = mod(psi[0].to(device), w=ang.to(device)).cpu().numpy()  # broadcast of a batch-1 state
    ref = fp64.apply_gate(psi.numpy().astype(np.complex128), N, ("ctrl", 2, ("ry", 0, ang.double().numpy())))
    assert np.abs(out - ref).max() < ATOL
    ref1 = fp64.apply_gate(np.broadcast_to(psi[0].numpy().astype(np.complex128), (B, 2 ** N)), N,
                           ("ctrl", 2, ("ry", 0, ang.double().numpy())))
    assert one.shape == (B, 2 ** N) and np.abs(one - ref1).max() < ATOL


def _loss_oracle(psi, ops, w):
    for op in ops:
        psi = fp64.apply_gate(psi, N, op)
    return float((fp64.measure_prob(psi, N) * w).sum())


@pytest.mark.parametrize("which", ["ctrl_ry", "ctrl_rx_named", "inv_rz", "u", "reset"])
def test_gradients_vs_finite_differences(device, which):
    B = 2
    gen = torch.Generator().manual_seed(11)
    psi0 = _state(N, B, seed=12)
    w = torch.randn(B, N, generator=gen)
    th = _f32(0.9)
    meas = qandle.MeasureProbability().build(num_qubits=N)
    named = {}
    if which == "ctrl_ry":
        mod = qandle.Controlled(1, qandle.RY(3, theta=th, remapping=None))
        op_of = lambda v: ("ctrl", 1, ("ry", 3, v))  # noqa: E731
    elif which == "ctrl_rx_named":
        mod = qandle.Controlled(4, qandle.RX(0, name="a", remapping=None))
        named = {"a": torch.tensor([0.3, -1.2])}
        op_of = lambda v: ("ctrl", 4, ("rx", 0, v))  # noqa: E731
    elif which == "inv_rz":
        mod = qandle.Invert(qandle.RZ(2, theta=th, remapping=None))
        op_of = lambda v: ("inv", ("rz", 2, v))  # noqa: E731
    elif which == "u":
        mod = qandle.U(3, torch.from_numpy(M1))
        op_of = lambda v: ("u", 3, M1.astype(np.complex128))  # noqa: E731
    else:
        mod = qandle.Reset(1)
        op_of = lambda v: ("reset", 1)  # noqa: E731
    with _nowarn():
        mod = mod.build(num_qubits=N).to(device)
    st = psi0.to(device).requires_grad_(True)
    kw = {k: v.to(device).requires_grad_(True) for k, v in named.items()}
    out = mod(st, **kw) if mod.named else mod(st)
    loss = (meas(out) * w.to(device)).sum()
    loss.backward()
    gpsi = st.grad.cpu().numpy()
    p0 = psi0.numpy().astype(np.complex128)
    wn = w.double().numpy()
    angle = named["a"].double().numpy() if named else th
    h = 1e-6
    # input-state gradient (torch convention dL/dRe + i dL/dIm) at a few amplitudes
    errs = []
    for b, i in [(0, 3), (1, 17), (0, 30)]:
        for part in (1.0, 1j):
            pp, pm = p0.copy(), p0.copy()
            pp[b, i] += h * part
            pm[b, i] -= h * part
            d = (_loss_oracle(pp, [op_of(angle)], wn) - _loss_oracle(pm, [op_of(angle)], wn)) / (2 * h)
            got = gpsi[b, i].real if part == 1.0 else gpsi[b, i].imag
            errs.append(abs(got - d))
    print(f"{which}: state-gradient max err {max(errs):.2e}")
    assert max(errs) < 1e-4
    if which in ("ctrl_ry", "inv_rz"):
        g = mod.t.theta.grad if which == "ctrl_ry" else mod.target.theta.grad
        d = (_loss_oracle(p0, [op_of(th + h)], wn) - _loss_oracle(p0, [op_of(th - h)], wn)) / (2 * h)
        assert abs(g.item() - d) < 1e-4, (g.item(), d)
    if which == "ctrl_rx_named":
        ga = kw["a"].grad.cpu().numpy()
        for b in range(B):
            ap, am = angle.copy(), angle.copy()
            ap[b] += h
            am[b] -= h
            d = (_loss_oracle(p0, [op_of(ap)], wn) - _loss_oracle(p0, [op_of(am)], wn)) / (2 * h)
            assert abs(ga[b] - d) < 1e-4, (b, ga[b], d)


def test_circuit_with_remaining_operators(device):
    """A circuit mixing every operator: lowered where the fused program covers
    it, operator by operator otherwise, against the oracle."""
    from qandle_amd._engine import fused

    n, B = 6, 3
    gen = torch.Generator().manual_seed(21)
    x = torch.rand(B, n, generator=gen) * 3
    a, b = _f32(0.45), _f32(2.1)
    layers = [qandle.AngleEmbedding(name="x", qubits=list(range(n)), rotation="ry")]
    layers += [qandle.RX(q, theta=_f32(0.2 + 0.3 * q), remapping=None) for q in range(n)]
    layers += [qandle.CZ(0, 5), qandle.SWAP(1, 4), qandle.Controlled(2, qandle.RY(3, theta=a, remapping=None)),
               qandle.U(0, torch.from_numpy(M1)), qandle.Invert(qandle.RZ(5, theta=b, remapping=None)),
               qandle.CNOT(3, 2), qandle.Reset(4), qandle.MeasureProbability()]
    circ = qandle.Circuit(layers, num_qubits=n).to(device)
    assert fused.lower_circuit(circ.circuit) is None  # Controlled, U, Reset: operator by operator
    with torch.no_grad():
        P = circ(x=x.to(device)).cpu().numpy()
    ops = [("embed", "y", list(range(n)), x.double().numpy())]
    ops += [("rx", q, _f32(0.2 + 0.3 * q)) for q in range(n)]
    ops += [("cz", 0, 5), ("swap", 1, 4), ("ctrl", 2, ("ry", 3, a)), ("u", 0, M1.astype(np.complex128)),
            ("inv", ("rz", 5, b)), ("cnot", 3, 2), ("reset", 4)]
    ref = fp64.measure_prob(fp64.simulate(n, ops), n)
    print(f"mixed circuit: max |P - fp64| = {np.abs(P - ref).max():.2e}")
    assert np.abs(P - ref).max() < ATOL


@pytest.mark.parametrize("specialise", [False, True])
def test_fused_swap_and_inverted_rotations(device, specialise):
    """SWAP (three CNOTs), Invert(CNOT) and Invert(rotation) (the negated angle)
    run inside the fused program; forward and adjoint gradients against fp64."""
    from qandle_amd import config
    from qandle_amd._engine import fused

    n, B = 7, 3
    gen = torch.Generator().manual_seed(31)
    x = torch.rand(B, n, generator=gen) * 3
    th = [_f32(0.3 + 0.4 * q) for q in range(n)]
    layers = [qandle.AngleEmbedding(name="x", qubits=list(range(n)), rotation="rx")]
    layers += [qandle.RY(q, theta=th[q], remapping=None) for q in range(n)]
    layers += [qandle.SWAP(0, 5), qandle.Invert(qandle.CNOT(2, 6)), qandle.Invert(qandle.SWAP(3, 1))]
    layers += [qandle.Invert(qandle.RX(q, theta=th[n - 1 - q], remapping=None)) for q in range(n)]
    layers += [qandle.CNOT(q, q + 1) for q in range(n - 1)]
    layers.append(qandle.MeasureProbability())
    old = config.SPECIALISE
    config.SPECIALISE = specialise
    try:
        with _nowarn():
            circ = qandle.Circuit(layers, num_qubits=n).to(device)
        assert fused.lower_circuit(circ.circuit) is not None
        wt = torch.randn(B, n, generator=gen)
        xx = x.to(device).requires_grad_(True)
        P = circ(x=xx)
        (P * wt.to(device)).sum().backward()
    finally:
        config.SPECIALISE = old
    ops = [("embed", "x", list(range(n)), x.double().numpy())]
    ops += [("ry", q, th[q]) for q in range(n)]
    ops += [("cnot", 5, 0), ("cnot", 0, 5), ("cnot", 5, 0), ("cnot", 2, 6), ("cnot", 1, 3), ("cnot", 3, 1), ("cnot", 1, 3)]
    ops += [("rx", q, -th[n - 1 - q]) for q in range(n)]
    ops += [("cnot", q, q + 1) for q in range(n - 1)]
    ref = fp64.measure_prob(fp64.simulate(n, ops), n)
    assert np.abs(P.detach().cpu().numpy() - ref).max() < ATOL
    adj = fp64.adjoint_grads(n, ops, wt.double().numpy())
    assert np.abs(xx.grad.cpu().numpy() - adj[0]).max() < 1e-4
    inv_mods = [m for m in circ.circuit.layers if isinstance(m, qandle.operators.BuiltInvert)
                and isinstance(m.target, qandle.operators.BuiltParametrizedOperator)]
    got = np.array([m.target.theta.grad.item() for m in inv_mods])
    # d/dtheta of RX(-theta) is minus the gradient of the gate's angle
    k0 = 1 + n + 7
    expect = -np.array([adj[k0 + q] for q in range(n)])
    assert np.abs(got - expect).max() < 1e-4
