"""QConv's host-side patch gathering and pixel scattering, on the CPU: the
circuit is replaced by a known map of each patch, so the test checks only the
unfold / row order / channel selection / rescale around it against an
explicit numpy loop over pixels (src/qandle/convolution.py:42-61)."""
import numpy as np
import pytest
import torch

import qandle_amd as qandle


class _FakeCircuit(torch.nn.Module):
    """A deterministic stand-in for the joint probabilities: position-weighted sums
    of the (padded) patch, distinct per output channel."""

    def __init__(self, q):
        super().__init__()
        self.q = q

    def forward(self, emb):
        q = self.q
        pad = torch.zeros(emb.shape[0], 2 ** q, dtype=emb.dtype)
        pad[:, : emb.shape[1]] = emb
        k = torch.arange(1, 2 ** q + 1, dtype=emb.dtype)
        return torch.stack([(pad * torch.roll(k, c)).sum(-1) for c in range(2 ** q)], dim=1)


@pytest.mark.parametrize("cin,cout,k,p,hw", [(2, 3, 3, 1, (5, 4)), (1, 5, (2, 3), (0, 1), (4, 6)), (3, 2, 1, 0, (3, 3))])
def test_patch_pipeline(cin, cout, k, p, hw):
    layer = qandle.QConv(in_channels=cin, out_channels=cout, kernel_size=k, padding=p)
    q = layer.qubits
    layer.qcircuit = _FakeCircuit(q)
    x = torch.rand(2, cin, *hw, generator=torch.Generator().manual_seed(1))
    out = layer(x).numpy()
    kh, kw = (k, k) if isinstance(k, int) else k
    ph, pw = (p, p) if isinstance(p, int) else p
    h, w = hw
    ho, wo = h + 2 * ph - kh + 1, w + 2 * pw - kw + 1
    xp = np.pad(x.numpy(), ((0, 0), (0, 0), (ph, ph), (pw, pw)))
    fake = _FakeCircuit(q)
    ref = np.zeros((2, cout, ho, wo))
    for bi in range(2):
        for i in range(ho):
            for j in range(wo):
                patch = xp[bi, :, i:i + kh, j:j + kw].reshape(-1) + 0.001  # channel, row, column order
                probs = fake(torch.from_numpy(patch[None]).float())[0].numpy()
                ref[bi, :, i, j] = probs[:cout] * cout / 2
    assert out.shape == (2, cout, ho, wo)
    np.testing.assert_allclose(out, ref, rtol=1e-5, atol=1e-5)


def test_register_size_and_errors():
    assert qandle.QConv(2, 3).qubits == 5           # 18 patch values -> 5 qubits
    assert qandle.QConv(1, 40, kernel_size=1, padding=0).qubits == 6  # 40 channels -> 6 qubits
    with pytest.raises(ZeroDivisionError):  # a one-qubit SEL has no CNOT ring, in the reference too
        qandle.QConv(1, 1, kernel_size=1, padding=0)
    layer = qandle.QConv(2, 3)
    with pytest.raises(ValueError, match="does not match in_channels"):
        layer(torch.zeros(1, 3, 4, 4))
    keys = list(layer.state_dict())
    assert keys and all(k.startswith("qcircuit.circuit.layers.1.mods.") and k.endswith(".theta") for k in keys)
