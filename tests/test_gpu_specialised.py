"""GPU parity of the run-time specialised pass kernels.

Each circuit runs with config.SPECIALISE on and off: the specialised kernels
must match the FP64 oracle to the fused-path bar (1e-5 on probabilities and
amplitudes) and the interpreted kernels to float32 rounding, including the
adjoint gradients that consume the specialised forward's states.
"""
import numpy as np
import pytest
import torch

import qandle_amd as qandle
from oracle import fp64
from qandle_amd import config, workloads

pytestmark = pytest.mark.gpu
ATOL = 1e-5


@pytest.fixture
def specialise():
    old = config.SPECIALISE
    yield
    config.SPECIALISE = old


def both(fn):
    config.SPECIALISE = False
    a = fn()
    config.SPECIALISE = True
    b = fn()
    return a, b


@pytest.mark.parametrize("n,B,depth", [(8, 32, 3), (13, 3, 2), (16, 2, 2), (20, 2, 10), (21, 2, 10)])
def test_specialised_ladder_vs_fp64(device, specialise, n, B, depth):
    w = workloads.rxryrz_ladder(n=n, B=B, depth=depth, seed=n + 1)
    circ = qandle.Circuit(w.layers, num_qubits=n).to(device)
    x = w.inputs["x"].to(device)

    def run():
        with torch.no_grad():
            return circ(x=x).cpu()

    P0, P1 = both(run)
    ref = fp64.measure_prob(fp64.simulate(n, w.oracle_ops()), n)
    err = np.abs(P1.numpy() - ref).max()
    print(f"specialised ladder n={n} B={B}: max |P - P_fp64| = {err:.2e}, vs interpreted {(P1 - P0).abs().max():.2e}")
    assert err < ATOL
    assert (P1 - P0).abs().max().item() < 1e-6


@pytest.mark.parametrize("n,B", [(10, 3), (11, 3), (12, 1), (13, 1), (13, 3), (14, 1), (14, 3), (21, 1)])
def test_default_direct_paths_at_tile_boundary_vs_fp64(device, specialise, n, B):
    """Widths around one tile per sample (a tile holds 2^11, 2^12 or 2^13
    amplitudes for the shapes chosen at 10-11, 12 and 13-21 qubits; 10 qubits
    packs two samples per tile), batches of 1 and 3: the default kernels, which
    load a pass's registers straight from the state when it fuses 32 or more
    unitaries and store them straight back where a wave writes whole lines,
    match the FP64 oracle and, bit for bit, the kernels that stage every
    register through the tile."""
    from qandle_amd._engine import fused

    w = workloads.rxryrz_ladder(n=n, B=B, depth=6, seed=300 + n + B)
    circ = qandle.Circuit(w.layers, num_qubits=n).to(device)
    x = w.inputs["x"].to(device)
    config.SPECIALISE = True
    old = config.SPECIALISE_OPTIONS
    try:
        config.SPECIALISE_OPTIONS = {}
        with torch.no_grad():
            P1 = circ(x=x).cpu()
        R, TB = fused.choose_shape(n)
        kern = fused.forward_kernels(fused.lower_circuit(circ.circuit), B, R, TB, device=device)
        config.SPECIALISE_OPTIONS = {"direct_load": False, "direct_store": False}
        with torch.no_grad():
            P0 = circ(x=x).cpu()
    finally:
        config.SPECIALISE_OPTIONS = old
    loads, stores = sum(a for a, _ in kern.direct), sum(b for _, b in kern.direct)
    ref = fp64.measure_prob(fp64.simulate(n, w.oracle_ops()), n)
    err = np.abs(P1.numpy() - ref).max()
    print(f"n={n} B={B}: {len(kern.direct)} passes, {loads} load and {stores} store registers directly; "
          f"max |P - P_fp64| = {err:.2e}")
    if n >= R + TB:
        assert stores >= 1 and (loads >= 1 or len(kern.direct) == 1)
    else:
        assert loads == stores == 0
    assert err < ATOL
    assert torch.equal(P0, P1)


@pytest.mark.parametrize("n,B", [(6, 5), (12, 4)])
def test_specialised_named_inputs_and_gradients(device, specialise, n, B):
    """Per-sample named rotations (the vector-ALU path in multi-sample tiles),
    a 0-d named input, uniform-control CNOTs; forward and adjoint gradients."""
    gen = torch.Generator().manual_seed(5 + n)
    x = torch.rand(B, n, generator=gen) * 3
    r = torch.rand(B, generator=gen) * 3
    s0 = torch.rand((), generator=gen) * 3
    layers = [qandle.AngleEmbedding(name="x", qubits=list(range(n)), rotation="ry")]
    layers += [qandle.RX(q, theta=0.4 + 0.3 * q, remapping=None) for q in range(n)]
    layers += [qandle.RX(1, name="r", remapping=None), qandle.CNOT(1, n - 2), qandle.RZ(n - 2, name="r", remapping=None),
               qandle.RY(0, name="s", remapping=None), qandle.CNOT(n - 1, 0), qandle.RZ(3, theta=0.7, remapping=None),
               qandle.MeasureProbability()]
    circ = qandle.Circuit(layers, num_qubits=n).to(device)
    wt = torch.randn(B, n, generator=gen).to(device)

    def run():
        xx = x.to(device).requires_grad_(True)
        rr = r.to(device).requires_grad_(True)
        ss = s0.to(device).requires_grad_(True)
        P = circ(x=xx, r=rr, s=ss)
        (P * wt).sum().backward()
        th = torch.stack([m.theta.grad.reshape(()) for m in circ.circuit.layers
                          if isinstance(m, qandle.operators.BuiltParametrizedOperator) and not m.named])
        for m in circ.circuit.layers:
            if isinstance(m, qandle.operators.BuiltParametrizedOperator) and not m.named:
                m.theta.grad = None
        return P.detach().cpu(), xx.grad.cpu(), rr.grad.cpu(), ss.grad.cpu(), th.cpu()

    a, b = both(run)
    for u, v, what in zip(a, b, ("P", "dx", "dr", "ds", "dtheta")):
        assert (u - v).abs().max().item() < 2e-5, what


def test_specialised_state_in_state_out(device, specialise):
    """No measurement, caller-provided input state (no product-state generator)."""
    gen = torch.Generator().manual_seed(9)
    n, B = 14, 3
    psi0 = torch.randn(B, 2 ** n, dtype=torch.cfloat, generator=gen)
    psi0 = psi0 / torch.linalg.norm(psi0, dim=-1, keepdim=True)
    layers = []
    for q in range(n):
        layers += [qandle.RY(q, theta=0.1 * q + 0.2, remapping=None), qandle.RZ(q, theta=0.3, remapping=None)]
    layers += [qandle.CNOT(q, (q + 3) % n) for q in range(n)]
    circ = qandle.Circuit(layers, num_qubits=n).to(device)
    st = psi0.to(device)

    def run():
        with torch.no_grad():
            return circ(st).cpu()

    s_a, s_b = both(run)
    assert (s_a - s_b).abs().max().item() < 1e-6
    ops = []
    for q in range(n):
        ops += [("ry", q, 0.1 * q + 0.2), ("rz", q, 0.3)]
    ops += [("cnot", q, (q + 3) % n) for q in range(n)]
    ops = [(k, q, float(np.float32(a))) if k != "cnot" else (k, q, a) for k, q, a in ops]
    ref = fp64.simulate(n, ops, psi0.numpy())
    assert np.abs(s_b.numpy() - ref).max() < ATOL


@pytest.mark.parametrize("n,B,depth", [(8, 5, 3), (12, 1, 3), (13, 3, 2), (14, 1, 4), (16, 2, 2)])
def test_specialised_adjoint_vs_fp64(device, specialise, n, B, depth):
    """Specialised forward + specialised adjoint sweep against the FP64 adjoint (1e-4)."""
    w = workloads.rxryrz_ladder(n=n, B=B, depth=depth, seed=200 + n, loss=True)
    circ = qandle.Circuit(w.layers, num_qubits=n).to(device)
    thetas = [m for m in circ.circuit.layers if isinstance(m, qandle.operators.BuiltParametrizedOperator)]

    def run():
        x = w.inputs["x"].to(device).requires_grad_(True)
        for m in thetas:
            m.theta.grad = None
        (circ(x=x) * w.loss_w.to(device)).sum().backward()
        return x.grad.cpu().numpy(), np.array([m.theta.grad.item() for m in thetas])

    (gx0, gt0), (gx1, gt1) = both(run)
    ops = w.oracle_ops()
    adj = fp64.adjoint_grads(n, ops, w.loss_w.double().numpy())
    ox = adj[0]
    ot = np.array([a for a, op in zip(adj[1:], ops[1:]) if op[0] != "cnot"])
    ex, et = np.abs(gx1 - ox).max(), np.abs(gt1 - ot).max()
    print(f"specialised adjoint n={n}: max|dx err|={ex:.2e} max|dtheta err|={et:.2e}, "
          f"vs interpreted {np.abs(gx1 - gx0).max():.2e} {np.abs(gt1 - gt0).max():.2e}")
    assert ex < 1e-4 and et < 1e-4
    assert np.abs(gx1 - gx0).max() < 1e-5 and np.abs(gt1 - gt0).max() < 1e-5


@pytest.mark.parametrize("n,B", [(7, 4), (15, 3)])
def test_specialised_folded_readout_and_gradients(device, specialise, n, B):
    """A random CNOT network ends the circuit: the compiled kernels fold it
    into parity readouts (forward) and parity co-state seeds (adjoint), with
    masks that include bit 0 (the pair bit of the seed's float4 loads)."""
    import math

    rng = np.random.default_rng(40 + n)
    body = []
    for q in range(n):
        body += [("ry", q, float(rng.uniform(0, 2 * math.pi))), ("rz", q, float(rng.uniform(0, 2 * math.pi)))]
    body += [("cnot", (q, q + 1)) for q in range(n - 1)]
    body += [("rx", q, float(rng.uniform(0, 2 * math.pi))) for q in range(n)]
    body += [("cnot", (n - 1, 0)), ("cnot", (0, n - 1)), ("cnot", (2, n - 1))]
    for _ in range(8):
        c, t = rng.choice(n, 2, replace=False)
        body.append(("cnot", (int(c), int(t))))
    x = torch.from_numpy(rng.uniform(0, math.pi, (B, n)).astype(np.float32))
    w = workloads._assemble(f"fold_n{n}", n, B, x, body)
    from qandle_amd._engine import fused

    prog = fused.program_from_spec(n, w.gates, w.embed)
    _, masks = fused.readout_fold(prog)
    assert any(m & 1 and m != 1 for m in masks)
    circ = qandle.Circuit(w.layers, num_qubits=n).to(device)
    thetas = [m for m in circ.circuit.layers if isinstance(m, qandle.operators.BuiltParametrizedOperator)]
    wt = torch.from_numpy(rng.standard_normal((B, n)).astype(np.float32))

    def run():
        xx = w.inputs["x"].to(device).requires_grad_(True)
        for m in thetas:
            m.theta.grad = None
        P = circ(x=xx)
        (P * wt.to(device)).sum().backward()
        return P.detach().cpu().numpy(), xx.grad.cpu().numpy(), np.array([m.theta.grad.item() for m in thetas])

    (P0, gx0, gt0), (P1, gx1, gt1) = both(run)
    ops = w.oracle_ops()
    ref = fp64.measure_prob(fp64.simulate(n, ops), n)
    adj = fp64.adjoint_grads(n, ops, wt.double().numpy())
    ot = np.array([a for a, op in zip(adj[1:], ops[1:]) if op[0] != "cnot"])
    print(f"folded readout n={n}: |P err| {np.abs(P1 - ref).max():.2e} |dx err| {np.abs(gx1 - adj[0]).max():.2e} "
          f"|dtheta err| {np.abs(gt1 - ot).max():.2e}")
    assert np.abs(P1 - ref).max() < ATOL and np.abs(P1 - P0).max() < 1e-6
    assert np.abs(gx1 - adj[0]).max() < 1e-4 and np.abs(gt1 - ot).max() < 1e-4
    assert np.abs(gx1 - gx0).max() < 1e-5 and np.abs(gt1 - gt0).max() < 1e-5


@pytest.mark.parametrize("opts", [{"pipeline": 8}, {"lanes": True}, {"lanes": False, "split_pairs": False},
                                  {"grid": "resident", "prefetch": True}, {"interleave": False},
                                  {"interleave": 2, "wave_sync": False}, {"stagger": 2},
                                  {"async_loads": True}, {"tail_prefetch": True}, {"direct_store": True},
                                  {"direct_load": True, "direct_store": True}])
def test_pass_kernel_variants_are_bit_identical(device, specialise, opts):
    """The pass-kernel variants (pipelined load/compute roles, operand loads,
    scalar operands with 16-byte tile moves, a resident grid with prefetch,
    one or two MFMA chains instead of four, block barriers between every
    sub-pass, a delayed start, direct global->LDS tile loads into two buffers,
    the next tile streamed in during the last sub-pass with direct stores,
    registers loaded from and stored to the state directly) compute exactly
    what the default kernels do."""
    from qandle_amd._engine import fused

    n, B = 16, 3
    w = workloads.rxryrz_ladder(n=n, B=B, depth=3, seed=17)
    circ = qandle.Circuit(w.layers, num_qubits=n).to(device)
    prog = fused.lower_circuit(circ.circuit)
    config.SPECIALISE = True
    old = config.SPECIALISE_OPTIONS
    try:
        with torch.no_grad():
            inp = fused.gather_inputs(prog, circ.circuit, None, {"x": w.inputs["x"].to(device)})
            cs = fused.build_coeffs(prog, inp)
            efac = fused.product_factors(prog, inp)
            R, TB = fused.choose_shape(n)
            config.SPECIALISE_OPTIONS = {}
            s0, p0 = fused.execute(prog, inp, cs, efac, R, TB, want_state=True)
            s0, p0 = s0.clone(), p0.clone()
            config.SPECIALISE_OPTIONS = opts
            s1, p1 = fused.execute(prog, inp, cs, efac, R, TB, want_state=True)
    finally:
        config.SPECIALISE_OPTIONS = old
    assert torch.equal(s0, s1) and torch.equal(p0, p1)
    ref = fp64.measure_prob(fp64.simulate(n, w.oracle_ops()), n)
    assert np.abs(p1.cpu().numpy() - ref).max() < ATOL


@pytest.mark.parametrize("embed", [True, False])
@pytest.mark.parametrize("n,B", [(9, 5), (14, 3)])
def test_leading_rotations_in_factors_vs_fp64(device, specialise, embed, n, B):
    """Forward without gradients from a product state (an AngleEmbedding, or
    |0..0>): the rotations ahead of every CNOT on their qubit, a per-sample
    named one among them, go into the qubits' factors (qs_factor_apply); the
    result equals the FP64 oracle and the same forward with them in the passes."""
    gen = torch.Generator().manual_seed(n * 7 + B)
    x = torch.rand(B, n, generator=gen) * 3
    r = torch.rand(B, generator=gen) * 3
    th = (torch.rand(3 * n + 4, generator=gen) * 6).tolist()
    layers = [qandle.AngleEmbedding(name="x", qubits=list(range(n)), rotation="ry")] if embed else []
    layers += [qandle.RX(q, theta=th[q], remapping=None) for q in range(n)]
    layers += [qandle.RZ(0, name="r", remapping=None), qandle.RY(n - 1, theta=th[n], remapping=None)]
    layers += [qandle.CNOT(q, q + 1) for q in range(n - 1)]
    layers += [qandle.RY(q, theta=th[n + 1 + q], remapping=None) for q in range(n)]
    layers += [qandle.RX(0, theta=th[2 * n + 1], remapping=None), qandle.CNOT(n - 1, 0), qandle.MeasureProbability()]
    circ = qandle.Circuit(layers, num_qubits=n).to(device)
    from qandle_amd._engine import fused

    prog = fused.lower_circuit(circ.circuit)
    lead = fused.leading_rotations(prog)
    assert lead is not None and sum(len(v) for v in lead[1].values()) == n + 2
    kw = {"r": r.to(device)}
    if embed:
        kw["x"] = x.to(device)
    old = config.ABSORB_LEADING_ROTATIONS
    try:
        outs = {}
        for absorb in (True, False):
            config.ABSORB_LEADING_ROTATIONS = absorb
            with torch.no_grad():
                outs[absorb] = both(lambda: circ(**kw).cpu().numpy())
    finally:
        config.ABSORB_LEADING_ROTATIONS = old
    ops = [("embed", "y", list(range(n)), x.double().numpy())] if embed else []
    ops += [("rx", q, th[q]) for q in range(n)]
    ops += [("rz", 0, r.double().numpy()), ("ry", n - 1, th[n])]
    ops += [("cnot", q, q + 1) for q in range(n - 1)]
    ops += [("ry", q, th[n + 1 + q]) for q in range(n)]
    ops += [("rx", 0, th[2 * n + 1]), ("cnot", n - 1, 0)]
    ref = fp64.measure_prob(fp64.simulate(n, ops), n)
    for absorb in (True, False):
        for P in outs[absorb]:
            assert np.abs(P - ref).max() < ATOL, (absorb, np.abs(P - ref).max())
    assert np.abs(outs[True][1] - outs[False][1]).max() < 1e-6


def test_wide_state_compiled_equals_interpreted(device, specialise):
    """26 qubits (a 1 GiB batch, tiles far from the state's top bits, lookahead
    plan with leading rotations in the factors): the compiled kernels agree with
    the interpreting ones, and the marginals lie in [0, 1]."""
    n, B = 26, 2
    w = workloads.rxryrz_ladder(n=n, B=B, depth=2, seed=3)
    circ = qandle.Circuit(w.layers, num_qubits=n).to(device)
    x = w.inputs["x"].to(device)

    def run():
        with torch.no_grad():
            return circ(x=x).cpu()

    p_interp, p_spec = both(run)
    assert p_spec.shape == (B, n)
    assert (p_spec - p_interp).abs().max().item() < 1e-6
    assert p_spec.min().item() >= -1e-6 and p_spec.max().item() <= 1 + 1e-6
