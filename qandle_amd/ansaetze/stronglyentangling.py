"""StronglyEntanglingLayer (``src/qandle/ansaetze/stronglyentangling.py:17-133``)."""
from __future__ import annotations

import copy
import typing

import torch

from .. import config
from .. import operators as op
from .. import qw_map
from .. import utils
from .ansatz import BuiltAnsatz, UnbuiltAnsatz

__all__ = ["StronglyEntanglingLayer"]


class StronglyEntanglingLayer(UnbuiltAnsatz):
    """Per layer: the rotations (default RZ, RY, RZ) on every qubit of ``qubits``,
    then a ring of CNOTs q_i -> q_(i + 1 + d mod (n_total - 1)) (``:17-78``).
    ``q_params`` (depth, len(qubits), len(rotations)) defaults to ``torch.rand``."""

    def __init__(
        self,
        qubits: typing.List[int],
        num_qubits_total: typing.Union[int, None] = None,
        depth: int = 1,
        rotations=("rz", "ry", "rz"),
        q_params=None,
        remapping: typing.Union[typing.Callable, None] = config.DEFAULT_MAPPING,
    ):
        super().__init__()
        self.depth = depth
        self.qubits = qubits
        self.num_qubits = num_qubits_total
        self.rots = [utils.parse_rot(r) for r in rotations]
        if q_params is None:
            q_params = torch.rand(depth, len(qubits), len(rotations))
        self.q_params = q_params
        if remapping is None:
            remapping = qw_map.none
        self.remapping = remapping

    def build(self, *args, **kwargs) -> BuiltAnsatz:
        return StronglyEntanglingLayerBuilt(
            num_qubits=kwargs["num_qubits"],
            depth=self.depth,
            rotations=self.rots,
            q_params=self.q_params,
            remapping=self.remapping,
            qubits=self.qubits,
        )

    def __str__(self) -> str:
        return "SEL"

    def to_qasm(self):
        raise NotImplementedError("TODO")  # the reference's unbuilt SEL has no export either (:57-58)

    def decompose(self) -> typing.List[op.UnbuiltOperator]:
        layers = []
        for d in range(self.depth):
            for wi, w in enumerate(self.qubits):
                for r in range(len(self.rots)):
                    layers.append(self.rots[r](qubit=w, theta=self.q_params[d, wi, r], remapping=self.remapping))
            layers.extend(StronglyEntanglingLayerBuilt._get_cnots(self.qubits, self.num_qubits, d % (self.num_qubits - 1)))
        return layers


class StronglyEntanglingLayerBuilt(BuiltAnsatz):
    """Built SEL (``:80-133``): a ``torch.nn.Sequential`` of built rotations and
    CNOTs (``mods``), so its ``state_dict`` keys are ``mods.<i>.theta``."""

    def __init__(self, num_qubits: int, qubits: typing.List[int], depth: int, rotations, q_params: torch.Tensor,
                 remapping: typing.Union[typing.Callable, None]):
        super().__init__()
        self.num_qubits = num_qubits
        self.depth = depth
        self.rots = rotations
        self.qubits = qubits
        layers = []
        for d in range(depth):
            for wi, w in enumerate(self.qubits):
                for r in range(len(self.rots)):
                    layers.append(self.rots[r](qubit=w, theta=q_params[d, wi, r], remapping=remapping).build(num_qubits))
            layers.extend(self._get_cnots(self.qubits, self.num_qubits, d % (len(self.qubits) - 1)))
        self.mods = torch.nn.Sequential(*layers)

    def forward(self, state: torch.Tensor):
        return self.mods(state)

    @staticmethod
    def _get_cnots(qubits: typing.List[int], num_qubits_total: int, iteration: int) -> typing.List[op.BuiltCNOT]:
        assert iteration + 1 < num_qubits_total
        cnots = []
        for ci in range(len(qubits)):
            ti = (ci + iteration + 1) % len(qubits)
            cnots.append(op.CNOT(qubits[ci], qubits[ti]).build(num_qubits_total))
        return cnots

    def __str__(self) -> str:
        return "SEL"

    def decompose(self) -> typing.List[op.Operator]:
        return [copy.deepcopy(m) for m in self.mods]
