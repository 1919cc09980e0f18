"""AngleEmbedding (``src/qandle/embeddings.py:114-214``).

The reference builds d dense (2^n x 2^n) rotation matrices per call and
multiplies them together (an O(B d 8^n) chain, ``:148-164``).  The product
state (x)_j R(x_j)|0> is separable, so the engine writes it directly:
``qs_embed_angle`` in the plain path, or as the generator of the first fused
pass inside a Circuit.
"""
from __future__ import annotations

import random
import typing

import torch

from . import operators as op
from . import qasm
from . import utils
from ._engine import plain

__all__ = ["AngleEmbedding", "AngleEmbeddingBuilt", "AmplitudeEmbedding", "AmplitudeEmbeddingBuilt",
           "InputOperator", "InputOperatorBuilt"]


class InputOperator(op.UnbuiltOperator):
    pass


class InputOperatorBuilt(op.BuiltOperator):
    def decompose(self):
        raise NotImplementedError(f"Decomposing {self.__class__} is not yet supported")

    def to_matrix(self, **kwargs):
        raise ValueError("Input operators do not have a matrix representation")


class AmplitudeEmbeddingBuilt(InputOperatorBuilt):
    """Ignores the incoming state and sets the features as the state
    (``src/qandle/embeddings.py:29-65``): ``kwargs[name]`` (d,) or (B, d) is
    padded with ``pad_with`` to 2^n entries (cropped when longer; without
    ``pad_with`` it must already have 2^n), L2-normalised when ``normalize``,
    and returned as real complex64 amplitudes (kernel ``qs_amplitude_embed``)."""

    def __init__(self, name: str, num_qubits: int, normalize: bool, pad_with: typing.Union[float, None]):
        super().__init__()
        self.num_qubits = num_qubits
        self.normalize = bool(normalize)
        self.pad_with = pad_with
        self.named = True
        self.name = name

    def embed(self, x: torch.Tensor) -> torch.Tensor:
        """(B, 2^n) state of (B, d) features (autograd-tracked)."""
        from ._engine.lib import require_device

        require_device(x, f"input '{self.name}'")
        N = 2 ** self.num_qubits
        if self.pad_with is None and x.shape[-1] != N:
            raise RuntimeError(f"AmplitudeEmbedding without pad_with needs {N} features, got {x.shape[-1]}")
        x2 = x.reshape(-1, x.shape[-1]).to(torch.float32)
        return plain.AmplitudeEmbedFn.apply(x2, self.num_qubits, 0.0 if self.pad_with is None else float(self.pad_with),
                                            self.normalize)

    def forward(self, _=None, **kwargs):
        x = kwargs[self.name]
        out = self.embed(x)
        return out.reshape(x.shape[:-1] + (2 ** self.num_qubits,))

    def __str__(self) -> str:
        return "AmplitudeEmbedding"

    def to_qasm(self) -> qasm.QasmRepresentation:
        raise NotImplementedError("AmplitudeEmbedding is not yet supported")


class AmplitudeEmbedding(InputOperator):
    """Amplitude embedding before build (``src/qandle/embeddings.py:68-111``)."""

    def __init__(self, name: str, qubits: list, normalize: bool = False, pad_with: typing.Union[float, None] = None):
        self.name = name
        self.named = True
        self.qubits = qubits
        self.pad_with = pad_with
        self.normalize = normalize

    def build(self, num_qubits: int, **kwargs) -> AmplitudeEmbeddingBuilt:
        assert num_qubits == len(self.qubits), "Current Implementation requires all qubits to be used."
        return AmplitudeEmbeddingBuilt(num_qubits=num_qubits, normalize=self.normalize, pad_with=self.pad_with,
                                       name=self.name)

    def __str__(self) -> str:
        return "AmplitudeEmbedding"

    def to_qasm(self) -> qasm.QasmRepresentation:
        raise NotImplementedError("AmplitudeEmbedding is not yet supported")


def _angle_qasm(rot, qubits):
    """One input-bound rotation per embedded qubit, named ``angle_<id>_<qubit>``
    with a fresh id per export (``src/qandle/embeddings.py:166-171, 206-211``)."""
    r = hash(random.random())
    return [rot(qubit=w, name=f"angle_{r}_{w}").to_qasm() for w in qubits]


_ROT_KIND = {"RX": 0, "RY": 1, "RZ": 2}


class AngleEmbeddingBuilt(InputOperatorBuilt):
    """Ignores the incoming state and returns (x)_q R(x_q / 2-angle)|0..0>.

    Input ``kwargs[name]`` has shape (d,) or (B, d); output (2^n,) or (B, 2^n)
    complex64.  The half-angle is x/2 with no remapping (``:150-151``).
    """

    def __init__(self, name: str, num_qubits: int, qubits: typing.List[int], rotation: str):
        super().__init__()
        self.num_qubits = num_qubits
        self.qubits = list(qubits)
        self.named = True
        self.name = name
        self.rotation = rotation
        self.rots = utils.parse_rot(rotation)
        self.kind = _ROT_KIND[self.rots.__name__]
        self.bits = [op.qubit_bit(num_qubits, q) for q in self.qubits]
        # The reference builds one throw-away rotation per embedded qubit
        # (``src/qandle/embeddings.py:140-146``); each consumes torch.rand(1) for
        # its unused theta.  Consume the same draws so that seeded parameter
        # initialisation of later layers matches the reference.
        for _ in self.qubits:
            torch.rand(1)

    def forward(self, state=None, **kwargs):
        x = kwargs[self.name]
        from ._engine.lib import require_device

        require_device(x, f"input '{self.name}'")
        if x.shape[-1] != len(self.qubits):
            raise RuntimeError(f"AngleEmbedding '{self.name}' expects {len(self.qubits)} features, got {x.shape[-1]}")
        if x.dim() == 1:
            return plain.EmbedFn.apply(x.unsqueeze(0), self.num_qubits, self.bits, self.kind).squeeze(0)
        if x.dim() == 2:
            return plain.EmbedFn.apply(x, self.num_qubits, self.bits, self.kind)
        lead = x.shape[:-1]
        out = plain.EmbedFn.apply(x.reshape(-1, x.shape[-1]), self.num_qubits, self.bits, self.kind)
        return out.reshape(lead + (2 ** self.num_qubits,))

    def __str__(self) -> str:
        return f"AngleEmbedding_{self.rotation}"

    def to_qasm(self) -> typing.List[qasm.QasmRepresentation]:
        return _angle_qasm(utils.parse_rot(self.rotation), self.qubits)

    def decompose(self):
        """The per-qubit named rotations this embedding is equivalent to."""
        return [self.rots(qubit=w, name=f"{self.name}_{w}") for w in self.qubits]


class AngleEmbedding(InputOperator):
    """Angle embedding (``src/qandle/embeddings.py:184-214``); ``name`` is required."""

    def __init__(self, name: str, qubits: typing.Union[typing.List[int], None] = None, rotation="rx"):
        self.name = name
        self.named = True
        self.qubits = qubits
        self.rotation = rotation

    def build(self, num_qubits: int, **kwargs) -> AngleEmbeddingBuilt:
        qubits = list(range(num_qubits)) if self.qubits is None else self.qubits
        return AngleEmbeddingBuilt(name=self.name, qubits=qubits, num_qubits=num_qubits, rotation=self.rotation)

    def __str__(self) -> str:
        return f"AngleEmbedding_{self.rotation}"

    def to_qasm(self) -> typing.List[qasm.QasmRepresentation]:
        return _angle_qasm(utils.parse_rot(self.rotation), self.qubits)
