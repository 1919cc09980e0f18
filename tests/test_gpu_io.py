"""GPU parity of AmplitudeEmbedding, MeasureJointProbability / MeasureState and
QConv (src/qandle/embeddings.py:29-111, measurements.py:73-91,
convolution.py:9-61) against the fp64 oracle; gradients against central
finite differences of the oracle (1e-4).
"""
import numpy as np
import pytest
import torch

import qandle_amd as qandle
from oracle import fp64
from qandle_amd._engine import fused

pytestmark = pytest.mark.gpu
ATOL = 1e-5
KIND = {qandle.operators.BuiltRX: "rx", qandle.operators.BuiltRY: "ry", qandle.operators.BuiltRZ: "rz"}


def _ops_of(mods):
    ops = []
    for m in mods:
        if isinstance(m, qandle.operators.BuiltCNOT):
            ops.append(("cnot", m.c, m.t))
        else:
            ops.append((KIND[type(m)], m.qubit, float(m.remapping(m.theta).reshape(-1)[0])))
    return ops


@pytest.mark.parametrize("d,n,pad,normalize", [(16, 4, None, True), (11, 4, 0.0, False), (11, 4, 0.25, True),
                                               (20, 4, 0.0, True), (5, 3, 1.5, False)])
def test_amplitude_embedding(device, d, n, pad, normalize):
    B = 3
    gen = torch.Generator().manual_seed(d + n)
    x = torch.randn(B, d, generator=gen)
    emb = qandle.AmplitudeEmbedding(name="a", qubits=list(range(n)), normalize=normalize, pad_with=pad).build(n)
    xx = x.to(device).requires_grad_(True)
    out = emb(a=xx)
    ref = fp64.amplitude_embed(x.double().numpy(), n, pad=pad, normalize=normalize)
    assert out.shape == (B, 2 ** n)
    assert np.abs(out.detach().cpu().numpy() - ref).max() < ATOL
    one = emb(a=x[0].to(device))
    assert one.shape == (2 ** n,) and np.abs(one.cpu().numpy() - ref[0]).max() < ATOL
    # gradient of L = sum c * Re(out) + e * Im(out) against finite differences
    c = torch.randn(B, 2 ** n, generator=gen)
    (out.real * c.to(device)).sum().backward()
    X = x.double().numpy()
    h = 1e-6
    for b, i in [(0, 0), (1, min(d - 1, 7)), (2, d - 1)]:
        xp, xm = X.copy(), X.copy()
        xp[b, i] += h
        xm[b, i] -= h
        fd = ((fp64.amplitude_embed(xp, n, pad, normalize).real - fp64.amplitude_embed(xm, n, pad, normalize).real)
              * c.double().numpy()).sum() / (2 * h)
        assert abs(xx.grad[b, i].item() - fd) < 1e-4, (b, i, xx.grad[b, i].item(), fd)


def test_joint_probability_and_state_readouts(device):
    n, B = 5, 4
    gen = torch.Generator().manual_seed(2)
    psi = torch.randn(B, 2 ** n, dtype=torch.complex64, generator=gen)
    st = psi.to(device).requires_grad_(True)
    P = qandle.MeasureJointProbability()(st)
    assert torch.equal(qandle.MeasureState()(st), st)
    np.testing.assert_allclose(P.detach().cpu().numpy(), fp64.joint_prob(psi.numpy().astype(np.complex128)),
                               rtol=0, atol=1e-6)
    w = torch.randn(B, 2 ** n, generator=gen)
    (P * w.to(device)).sum().backward()
    np.testing.assert_allclose(st.grad.cpu().numpy(), (2 * w * psi).numpy(), rtol=0, atol=1e-6)


@pytest.mark.parametrize("specialise", [False, True])
def test_amplitude_sel_joint_circuit(device, specialise):
    """AmplitudeEmbedding -> SEL -> MeasureJointProbability as one fused program."""
    from qandle_amd import config

    n, B, depth = 6, 5, 3
    gen = torch.Generator().manual_seed(8)
    x = torch.randn(B, 40, generator=gen)
    weights = torch.rand(depth, n, 3, generator=gen) * 6
    circ = qandle.Circuit([qandle.AmplitudeEmbedding(name="e", qubits=list(range(n)), pad_with=0.1, normalize=True),
                           qandle.StronglyEntanglingLayer(qubits=list(range(n)), depth=depth, q_params=weights,
                                                          remapping=None),
                           qandle.MeasureJointProbability()], num_qubits=n).to(device)
    prog = fused.lower_circuit(circ.circuit)
    assert prog is not None and prog.init[0] == "amplitude" and prog.post == "abs2"
    old = config.SPECIALISE
    config.SPECIALISE = specialise
    try:
        xx = x.to(device).requires_grad_(True)
        P = circ(e=xx)
        w = torch.randn(B, 2 ** n, generator=gen)
        (P * w.to(device)).sum().backward()
    finally:
        config.SPECIALISE = old
    ops = _ops_of(circ.circuit.layers[1].mods)
    X, W = x.double().numpy(), w.double().numpy()

    def loss(Xv, ops_):
        psi = fp64.simulate(n, ops_, fp64.amplitude_embed(Xv, n, pad=0.1, normalize=True))
        return float((fp64.joint_prob(psi) * W).sum()), fp64.joint_prob(psi)

    _, ref = loss(X, ops)
    assert np.abs(P.detach().cpu().numpy() - ref).max() < ATOL
    h = 1e-6
    for b, i in [(0, 3), (4, 39)]:
        xp, xm = X.copy(), X.copy()
        xp[b, i] += h
        xm[b, i] -= h
        fd = (loss(xp, ops)[0] - loss(xm, ops)[0]) / (2 * h)
        assert abs(xx.grad[b, i].item() - fd) < 1e-4, (b, i, xx.grad[b, i].item(), fd)
    thetas = [m for m in circ.circuit.layers[1].mods if hasattr(m, "theta")]
    for k in (0, len(thetas) // 2, len(thetas) - 1):
        oi = [j for j, op in enumerate(ops) if op[0] != "cnot"][k]
        op_p, op_m = list(ops), list(ops)
        op_p[oi] = (ops[oi][0], ops[oi][1], ops[oi][2] + h)
        op_m[oi] = (ops[oi][0], ops[oi][1], ops[oi][2] - h)
        fd = (loss(X, op_p)[0] - loss(X, op_m)[0]) / (2 * h)
        assert abs(thetas[k].theta.grad.item() - fd) < 1e-4, (k, thetas[k].theta.grad.item(), fd)


def _unfold(img, kh, kw, ph, pw):
    """im2col of (b, c, h, w) with zero padding: (b, c * kh * kw, h_out * w_out), torch's Unfold order."""
    b, c, h, w = img.shape
    p = np.zeros((b, c, h + 2 * ph, w + 2 * pw))
    p[:, :, ph:ph + h, pw:pw + w] = img
    ho, wo = h + 2 * ph - kh + 1, w + 2 * pw - kw + 1
    cols = np.zeros((b, c * kh * kw, ho * wo))
    for ci in range(c):
        for i in range(kh):
            for j in range(kw):
                row = (ci * kh + i) * kw + j
                cols[:, row, :] = p[:, ci, i:i + ho, j:j + wo].reshape(b, -1)
    return cols, ho, wo


def test_qconv(device):
    torch.manual_seed(11)
    layer = qandle.QConv(in_channels=2, out_channels=3, kernel_size=3, padding=1, qdepth=2).to(device)
    gen = torch.Generator().manual_seed(12)
    img = torch.rand(2, 2, 5, 4, generator=gen)
    xx = img.to(device).requires_grad_(True)
    out = layer(xx)
    assert out.shape == (2, 3, 5, 4)
    q = layer.qubits
    ops = _ops_of(layer.qcircuit.circuit.layers[1].mods)

    def forward(I):
        cols, ho, wo = _unfold(I, 3, 3, 1, 1)
        feats = cols.transpose(0, 2, 1).reshape(-1, cols.shape[1]) + 0.001
        P = fp64.joint_prob(fp64.simulate(q, ops, fp64.amplitude_embed(feats, q, pad=0, normalize=False)))
        P = P.reshape(I.shape[0], ho, wo, -1).transpose(0, 3, 1, 2)
        return P[:, :3] * 3 / 2

    ref = forward(img.double().numpy())
    assert np.abs(out.detach().cpu().numpy() - ref).max() < ATOL
    w = torch.randn(out.shape, generator=gen)
    (out * w.to(device)).sum().backward()
    I0, W = img.double().numpy(), w.double().numpy()
    h = 1e-6
    for idx in [(0, 0, 0, 0), (1, 1, 2, 3), (0, 1, 4, 1)]:
        ip, im = I0.copy(), I0.copy()
        ip[idx] += h
        im[idx] -= h
        fd = ((forward(ip) - forward(im)) * W).sum() / (2 * h)
        assert abs(xx.grad[idx].item() - fd) < 1e-4, (idx, xx.grad[idx].item(), fd)
