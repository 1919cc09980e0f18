"""GPU parity at the benchmark workloads' full sizes, against FP64 fixtures.

The FP64 values come from tests/golden/make_config_fixtures.py (run in the
build container; the GPU machine only compares).  Each test first checks that the
seeded generators still produce the circuit the fixture was computed for.

Bars: probabilities 1e-5 absolute, gradients 1e-4 absolute (the FP64 error of
each workload is printed); results that must not depend on scheduling (a
serving request alone vs in a mix, chunked vs whole) are compared bit for bit.
"""
import math
import os

import numpy as np
import pytest
import scipy.linalg
import torch

import qandle_amd as qandle
from qandle_amd import config, workloads
from qandle_amd._engine import fused, serving
from qandle_amd._engine.uneven import UnevenBatch
from tests.golden import make_config_fixtures as FX

pytestmark = pytest.mark.gpu
ATOL, GTOL = 1e-5, 1e-4
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name))


def _thetas(w):
    return np.array([g[2] for g in w.gates if g[0] != "cnot"], dtype=np.float32)


def _same_circuit(w, fx_x, fx_theta):
    assert np.array_equal(w.inputs["x"].numpy(), fx_x), "seeded inputs changed since the fixture was written"
    assert np.array_equal(_thetas(w), fx_theta), "seeded angles changed since the fixture was written"


def test_c2_full_batch(device):
    """20 qubits x 64, 10 layers: four rows of the batch-64 run against FP64."""
    fx = _load("c2_rows.npz")
    w = workloads.rxryrz_ladder(n=20, B=64, depth=10, seed=1)
    _same_circuit(w, fx["x"], fx["theta"])
    circ = qandle.Circuit(w.layers, num_qubits=20).to(device)
    with torch.no_grad():
        P = circ(x=w.inputs["x"].to(device)).cpu().numpy()
    assert P.shape == (64, 20) and np.all(np.isfinite(P))
    err = np.abs(P[fx["rows"]] - fx["P"]).max()
    print(f"C2 n=20 B=64: max |P - P_fp64| over rows {list(fx['rows'])} = {err:.2e}")
    assert err < ATOL


def test_c3_full_batch_gradients(device):
    """16 qubits x 256, 20 layers, checkpoint every 5 layers: dL/dtheta summed over all
    256 samples (960 angles) and dL/dx, against the FP64 adjoint."""
    fx = _load("c3_grads.npz")
    w = workloads.rxryrz_ladder(n=16, B=256, depth=20, seed=2, loss=True)
    _same_circuit(w, fx["x"], fx["theta"])
    assert np.array_equal(w.loss_w.numpy(), fx["w"])
    circ = qandle.Circuit(w.layers, num_qubits=16).to(device)
    x = w.inputs["x"].to(device).requires_grad_(True)
    P = circ(x=x)
    (P * w.loss_w.to(device)).sum().backward()
    thetas = [m for m in circ.circuit.layers if isinstance(m, qandle.operators.BuiltParametrizedOperator)]
    gth = np.array([m.theta.grad.item() for m in thetas])
    ep = np.abs(P.detach().cpu().numpy() - fx["P"]).max()
    ex = np.abs(x.grad.cpu().numpy() - fx["grad_x"]).max()
    eth = np.abs(gth - fx["grad_theta"]).max()
    print(f"C3 n=16 B=256: max|P err|={ep:.2e} max|dx err|={ex:.2e} max|dtheta err|={eth:.2e} "
          f"(|dtheta| up to {np.abs(fx['grad_theta']).max():.2f})")
    assert ep < ATOL and ex < GTOL and eth < GTOL


@pytest.mark.timeout(900)
@pytest.mark.parametrize("R,TB", [(4, 8), (4, 7)])
def test_c4_full_mix(device, R, TB):
    """All 2048 circuits (4-18 qubits, depth 2-24) in one batch, forward + adjoint in
    one captured graph, replayed twice: every circuit against its own single-circuit
    run, and the FP64 subset (every 17-18-qubit circuit of depth >= 20 and a spread
    of others) against the oracle."""
    fx = _load("c4_subset.npz")
    ws = workloads.random_uneven(2048)
    gen = torch.Generator().manual_seed(3)
    lw = [torch.randn(w.B, w.n, generator=gen) for w in ws]
    sel = [int(i) for i in fx["index"]]
    assert sel == FX.c4_selection(ws)
    for i in sel:
        assert np.array_equal(ws[i].inputs["x"].numpy()[0], fx[f"x{i}"])
        assert np.array_equal(_thetas(ws[i]), fx[f"theta{i}"])
        assert np.allclose(lw[i].double().numpy()[0], fx[f"w{i}"])
    ub = UnevenBatch.from_workloads(ws, device, R=R, TB=TB, cost=fused.UNEVEN_COST)
    ub.set_loss_weights(lw)
    ub.capture(backward=True)
    for rep in range(2):
        ub.replay(backward=True)
        torch.cuda.synchronize()
        _ = torch.ones(3, device=device).sum().item()  # eager work between replays
    # FP64 subset
    ep = eg = 0.0
    for i in sel:
        ep = max(ep, np.abs(ub.probs(i).cpu().numpy().reshape(-1) - fx[f"P{i}"]).max())
        eg = max(eg, np.abs(ub.input_grads(i).cpu().numpy().reshape(-1) - fx[f"gx{i}"]).max())
        eg = max(eg, np.abs(ub.param_grads(i).cpu().numpy() - fx[f"gth{i}"]).max())
    print(f"C4 R={R} TB={TB}: {len(sel)} circuits vs FP64: max|P err|={ep:.2e} max|grad err|={eg:.2e}")
    assert ep < ATOL and eg < GTOL
    # every circuit against the single-circuit path (library kernels, no per-circuit compile)
    old = config.SPECIALISE
    config.SPECIALISE = False
    try:
        dp = dg = 0.0
        for i, w in enumerate(ws):
            circ = qandle.Circuit(w.layers, num_qubits=w.n).to(device)
            x = w.inputs["x"].to(device).requires_grad_(True)
            P = circ(x=x)
            (P * lw[i].to(device)).sum().backward()
            th = torch.stack([m.theta.grad.reshape(()) for m in circ.circuit.layers
                              if isinstance(m, qandle.operators.BuiltParametrizedOperator)])
            dp = max(dp, (ub.probs(i).reshape(P.shape) - P.detach()).abs().max().item())
            dg = max(dg, (ub.input_grads(i).reshape(x.shape) - x.grad).abs().max().item(),
                     (ub.param_grads(i) - th).abs().max().item())
    finally:
        config.SPECIALISE = old
    print(f"C4 R={R} TB={TB}: 2048 circuits vs single-circuit runs: max|P diff|={dp:.2e} max|grad diff|={dg:.2e}")
    assert dp < ATOL and dg < GTOL


def _c5_requests(device, arrivals=None):
    reqs = []
    for k, (n, B, depth, seed) in enumerate(FX.C5_REQUESTS):
        w = FX.c5_request(n, B, depth, seed)
        circ = qandle.Circuit(w.layers, num_qubits=n).to(device)
        t = 0.0 if arrivals is None else arrivals[k]
        reqs.append((serving.Request(k, circ, {"x": w.inputs["x"].to(device)}, arrival=t), w))
    return reqs


def test_c5_serving_full_sizes(device):
    """Requests of 16-24 qubits and batches up to 128 under a 1 GiB cap: the 22- and
    24-qubit requests are split into several batch chunks.  Each result must be
    bit-identical served alone, in the mix, and whole (no cap); requests of at
    most 20 qubits are checked against FP64."""
    fx = _load("c5_rows.npz")
    cap = 2 ** 30
    mix = _c5_requests(device, arrivals=[0.0, 0.001, 0.002, 0.003, 0.004, 0.005])
    for k, (r, w) in enumerate(mix):
        _same_circuit(w, fx[f"x{k}"], fx[f"theta{k}"])
    eng = serving.ServingEngine(device, cap_bytes=cap, streams=4)
    for r, _ in mix:
        eng.prepare(r)
    chunks = {r.rid: len(r.chunks) for r, _ in mix}
    assert chunks[3] >= 2 and chunks[4] >= 2, chunks
    done = eng.serve([r for r, _ in mix])
    assert len(done) == len(mix) and all(r.error is None for r in done)
    assert eng.peak_resident <= cap
    alone = _c5_requests(device)
    whole = _c5_requests(device)
    for (a, _), (m, _), (h, _) in zip(alone, mix, whole):
        solo = serving.ServingEngine(device, cap_bytes=cap, streams=1, max_inflight=1)
        solo.prepare(a)
        solo.serve([a])
        big = serving.ServingEngine(device, cap_bytes=200e9, streams=1, max_inflight=1)
        big.prepare(h)
        assert len(h.chunks) == 1
        big.serve([h])
        assert torch.equal(a.result, m.result), f"request {m.rid}: alone and mixed results differ"
        assert torch.equal(h.result, m.result), f"request {m.rid}: chunked and whole results differ"
        assert m.result.shape == (m.B, m.n)
    err = 0.0
    for k, (r, w) in enumerate(mix):
        if f"P{k}" in fx:
            err = max(err, np.abs(r.result.cpu().numpy()[fx[f"rows{k}"]] - fx[f"P{k}"]).max())
    print(f"C5 requests <= 20 qubits: max |P - P_fp64| = {err:.2e}; chunks {chunks}")
    assert err < ATOL
    st = serving.latency_stats(done)
    assert st["p99_ms"] >= st["p50_ms"] > 0


def _closed_form_marginals(x, layers_u):
    """P(measured bit q = 0) for an RX product state, per-qubit 2x2 unitaries, then a
    CNOT ladder 0->1->...->n-1: the ladder writes the parity of qubits 0..q into
    qubit q, and for independent qubits P(parity = 0) = (1 + prod_j (p0_j - p1_j)) / 2."""
    B, n = x.shape
    out = np.empty((B, n))
    for b in range(B):
        z = []
        for q in range(n):
            v = scipy.linalg.expm(-0.5j * x[b, q] * np.array([[0, 1], [1, 0]])) @ np.array([1.0, 0.0])
            for U in layers_u[q]:
                v = U @ v
            z.append(abs(v[0]) ** 2 - abs(v[1]) ** 2)
        out[b] = (1.0 + np.cumprod(z)) / 2.0
    return out


@pytest.mark.timeout(600)
def test_28_qubits_closed_form(device):
    """28 qubits x 4 (8 GiB of state): RX embedding, two layers of RX, RY, RZ per qubit,
    a CNOT ladder, MeasureProbability.  Closed-form marginals in FP64 (scipy expm of
    the Pauli generators); checked with the leading rotations on the qubit factors
    (default), with every rotation applied to the full state in the pass kernels,
    and through the plain per-gate path; the unmeasured state has unit norm."""
    n, B = 28, 4
    g = torch.Generator().manual_seed(28)
    x = (torch.rand(B, n, generator=g) * math.pi).numpy().astype(np.float32)
    th = (torch.rand(2, n, 3, generator=g) * 2 * math.pi).numpy().astype(np.float32)
    P_ = {"x": np.array([[0, 1], [1, 0]]), "y": np.array([[0, -1j], [1j, 0]]), "z": np.diag([1.0, -1.0])}
    layers_u = [[scipy.linalg.expm(-0.5j * float(th[l, q, k]) * P_[r]) for l in range(2) for k, r in enumerate("xyz")]
                for q in range(n)]
    ref = _closed_form_marginals(x.astype(np.float64), layers_u)
    body = [qandle.AngleEmbedding(name="x", qubits=list(range(n)))]
    ctor = {"x": qandle.RX, "y": qandle.RY, "z": qandle.RZ}
    for l in range(2):
        for q in range(n):
            for k, r in enumerate("xyz"):
                body.append(ctor[r](q, theta=float(th[l, q, k]), remapping=None))
    body += [qandle.CNOT(q, q + 1) for q in range(n - 1)]
    circ = qandle.Circuit(body + [qandle.MeasureProbability()], num_qubits=n).to(device)
    xd = torch.from_numpy(x).to(device)
    errs = {}
    with torch.no_grad():
        errs["fused"] = np.abs(circ(x=xd).cpu().numpy() - ref).max()
        old = config.ABSORB_LEADING_ROTATIONS
        config.ABSORB_LEADING_ROTATIONS = False
        try:
            fresh = qandle.Circuit(body + [qandle.MeasureProbability()], num_qubits=n).to(device)
            errs["fused, rotations on the state"] = np.abs(fresh(x=xd).cpu().numpy() - ref).max()
        finally:
            config.ABSORB_LEADING_ROTATIONS = old
        errs["plain per-gate"] = np.abs(circ.circuit.forward_layers(x=xd).cpu().numpy() - ref).max()
        state_circ = qandle.Circuit(body, num_qubits=n).to(device)
        psi = state_circ(x=xd)
        # float64 sums: a float32 sum of 2^28 terms drifts by about 1e-4 by itself
        a2 = psi.abs().to(torch.float64) ** 2
        norm = a2.sum(-1).cpu().numpy()
        p0 = a2.reshape(B, 2, -1)[:, 0].sum(-1).cpu().numpy()  # qubit 0 is the top bit
        del a2
        del psi
    print(f"28 qubits x 4: max |P - closed form| = {errs}; norm - 1 = {np.abs(norm - 1).max():.2e}")
    assert all(e < ATOL for e in errs.values()), errs
    assert np.abs(norm - 1).max() < 1e-4
    assert np.abs(p0 - ref[:, 0]).max() < ATOL  # qubit 0's marginal is untouched by the ladder
