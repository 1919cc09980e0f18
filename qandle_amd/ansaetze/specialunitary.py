"""SpecialUnitary / SU (``src/qandle/ansaetze/specialunitary.py:11-123``)."""
from __future__ import annotations

import typing

import torch

from .. import operators as op
from .. import utils
from .ansatz import BuiltAnsatz, UnbuiltAnsatz
from .twolocal import column_cnot_chain

__all__ = ["SpecialUnitary"]


class SpecialUnitary(UnbuiltAnsatz):
    """Hardware-efficient SU(2) 2-local circuit (``:11-56``)."""

    def __init__(self, qubits: typing.Union[typing.List[int], None] = None, reps: int = 1, rotations=("ry",)):
        super().__init__()
        self.qubits = qubits
        self.reps = reps
        self.rotations = [utils.parse_rot(r) for r in rotations]

    def build(self, *args, **kwargs) -> BuiltAnsatz:
        return SpecialUnitaryBuilt(num_qubits=kwargs["num_qubits"], reps=self.reps, rotations=self.rotations,
                                   qubits=self.qubits)

    def __str__(self) -> str:
        return "SU"

    def decompose(self) -> typing.List[op.Operator]:
        if self.qubits is None:
            raise ValueError(
                "It is not specified on which qubits to apply the ansatz. Build the circuit to specify or pass "
                "them explicitly as a list"
            )
        outp = []
        for r in self.rotations:
            for w in self.qubits:
                outp.append(r(w))
        for _ in range(self.reps):
            for wi in range(len(self.qubits) - 1):
                outp.append(op.CNOT(self.qubits[wi], self.qubits[wi + 1]))
            for r in self.rotations:
                for w in self.qubits:
                    outp.append(r(w))
        return outp


class SpecialUnitaryBuilt(BuiltAnsatz):
    """Built SU (``:59-123``): reps + 1 rotation blocks (default-initialised
    rotations, one ``torch.rand(1)`` each at build) with the CNOT chain applied
    between them as ``cnots @ state`` (see :func:`twolocal.column_cnot_chain`)."""

    def __init__(self, num_qubits: int, qubits: typing.Union[typing.List[int], None], reps: int, rotations):
        super().__init__()
        self.num_qubits = num_qubits
        self.reps = reps
        self.rotations = rotations
        self.qubits = qubits or list(range(num_qubits))
        layers = []
        for _ in range(reps + 1):
            block = []
            for w in self.qubits:
                for r in rotations:
                    block.append(r(w))
            layers.append(block)
        self.layers = torch.nn.ModuleList(
            [torch.nn.Sequential(*[la.build(num_qubits=num_qubits) for la in lay]) for lay in layers])

    def forward(self, state: torch.Tensor) -> torch.Tensor:
        state = self.layers[0](state)
        for layer in self.layers[1:]:
            state = column_cnot_chain(state, self.qubits, self.num_qubits)
            state = layer(state)
        return state

    def __str__(self) -> str:
        return "SU"

    def decompose(self) -> typing.List[op.Operator]:
        res = []
        res.extend(self.layers[0])
        for mod in self.layers[1:]:
            res.extend(mod)
            for w in range(len(self.qubits) - 1):
                res.append(op.CNOT(self.qubits[w], self.qubits[w + 1]))
        return res
