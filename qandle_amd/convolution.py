"""QConv (``src/qandle/convolution.py:9-61``): a quantum convolution layer.

Each output pixel is one small circuit: the input patch under the kernel (all
input channels, ``in_channels * kh * kw`` values, shifted by +0.001 so no
patch is all zero) is amplitude-embedded into ``q`` qubits (zero-padded to
2^q), a StronglyEntanglingLayer of depth ``qdepth`` acts on it and the joint
basis-state probabilities are read out.  The first ``out_channels``
probabilities, times ``out_channels / 2``, are the pixel's output channels.

Engine mapping: the patches of the whole batch are the batch of ONE circuit
call, so a layer forward is one fused program over ``b * h_out * w_out``
samples (amplitude-embedding kernel, the SEL's passes, |psi|^2 readout)
instead of one small circuit per pixel.  The host side here only gathers
patches and scatters pixels, with plain tensor reshapes.
"""
from __future__ import annotations

import math
import typing

import torch

from . import embeddings, measurements, qcircuit
from .ansaetze import StronglyEntanglingLayer

__all__ = ["QConv"]

PATCH_SHIFT = 0.001  # added to every patch value before embedding (a zero patch has no direction)


def _pair(v) -> typing.Tuple[int, int]:
    return v if isinstance(v, tuple) else (v, v)


def register_size(in_channels: int, out_channels: int, kernel: typing.Tuple[int, int]) -> int:
    """Qubits of the patch circuit: enough amplitudes for a patch and enough
    basis states for the output channels, at least one."""
    need_in = math.ceil(math.log2(kernel[0] * kernel[1] * in_channels))
    need_out = math.ceil(math.log2(out_channels))
    return max(need_in, need_out, 1)


class QConv(torch.nn.Module):
    """Same constructor, attributes ``qubits``, ``unfold`` and ``qcircuit`` (so
    ``state_dict`` keys ``qcircuit.circuit.layers.1.mods.<i>.theta``) and error
    text as the reference."""

    def __init__(self, in_channels: int, out_channels: int,
                 kernel_size: typing.Union[typing.Tuple[int, int], int] = (3, 3),
                 padding: typing.Union[typing.Tuple[int, int], int] = (1, 1), qdepth: int = 1):
        super().__init__()
        self.in_channels, self.out_channels, self.qdepth = in_channels, out_channels, qdepth
        self.kernel_size, self.padding = _pair(kernel_size), _pair(padding)
        self.unfold = torch.nn.Unfold(kernel_size=kernel_size, padding=self.padding)
        self.qubits = register_size(in_channels, out_channels, self.kernel_size)
        wires = list(range(self.qubits))
        self.qcircuit = qcircuit.Circuit(
            num_qubits=self.qubits,
            layers=[embeddings.AmplitudeEmbedding(qubits=wires, pad_with=0, name="emb"),
                    StronglyEntanglingLayer(qubits=wires, depth=qdepth),
                    measurements.MeasureJointProbability()],
        )

    def output_size(self, h: int, w: int) -> typing.Tuple[int, int]:
        (kh, kw), (ph, pw) = self.kernel_size, self.padding
        return h + 2 * ph - kh + 1, w + 2 * pw - kw + 1

    def forward(self, x):
        b, c_in, h, w = x.shape
        if c_in != self.in_channels:
            raise ValueError(f"Input channels {c_in} does not match in_channels {self.in_channels}")
        h_out, w_out = self.output_size(h, w)
        cols = self.unfold(x)                                  # (b, c_in*kh*kw, h_out*w_out)
        patches = cols.transpose(1, 2).reshape(-1, cols.shape[1]) + PATCH_SHIFT  # pixel-major rows
        probs = self.qcircuit(emb=patches)                     # (b*h_out*w_out, 2^q)
        img = probs.reshape(b, h_out, w_out, -1).permute(0, 3, 1, 2)
        return img[:, : self.out_channels] * (self.out_channels / 2)
