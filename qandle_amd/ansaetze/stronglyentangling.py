"""StronglyEntanglingLayer (``src/qandle/ansaetze/stronglyentangling.py:17-133``).

The gate pattern comes from :func:`schedules.strongly_entangling`; this module
only materialises it.  Public surface kept from the reference: the
constructor arguments, the construction-time ``torch.rand(depth, len(qubits),
len(rotations))`` draw of ``q_params``, ``StronglyEntanglingLayerBuilt.mods``
(so ``state_dict`` keys are ``mods.<i>.theta``), ``decompose`` and the
error types.
"""
from __future__ import annotations

import copy
import typing

import torch

from .. import config
from .. import operators as op
from .. import qw_map
from .. import utils
from . import schedules
from .ansatz import BuiltAnsatz, UnbuiltAnsatz

__all__ = ["StronglyEntanglingLayer"]


def _unbuilt(spec, q_params, remapping):
    """Unbuilt operator of one schedule record (rotation angles from ``q_params``)."""
    if isinstance(spec, schedules.CtrlSpec):
        return spec.gate(spec.control, spec.target)
    return spec.rot(qubit=spec.qubit, theta=q_params[spec.param], remapping=remapping)


class StronglyEntanglingLayer(UnbuiltAnsatz):
    """Per layer: each of ``rotations`` (default RZ, RY, RZ) on every qubit of
    ``qubits``, then a CNOT ring whose offset grows by one per layer."""

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
        self.qubits, self.num_qubits, self.depth = qubits, num_qubits_total, depth
        self.rots = [utils.parse_rot(r) for r in rotations]
        # drawn here, at construction (the built layer only reads it)
        self.q_params = torch.rand(depth, len(qubits), len(rotations)) if q_params is None else q_params
        self.remapping = qw_map.none if remapping is None else remapping

    def _schedule(self, period, total):
        return schedules.strongly_entangling(self.qubits, self.depth, self.rots, period, total, op.CNOT)

    def build(self, *args, **kwargs) -> BuiltAnsatz:
        return StronglyEntanglingLayerBuilt(num_qubits=kwargs["num_qubits"], qubits=self.qubits, depth=self.depth,
                                            rotations=self.rots, q_params=self.q_params, remapping=self.remapping)

    def __str__(self) -> str:
        return "SEL"

    def to_qasm(self):
        # the reference has no statement form for an unbuilt layer either (``:57-58``)
        raise NotImplementedError("an unbuilt StronglyEntanglingLayer has no OpenQASM form; build it first")

    def decompose(self) -> typing.List[op.Operator]:
        """Unbuilt rotations and built CNOTs, as the reference returns them
        (``:60-75``); the ring offset here runs modulo ``num_qubits_total - 1``."""
        total = self.num_qubits
        out = []
        for spec in self._schedule(total - 1, total):
            g = _unbuilt(spec, self.q_params, self.remapping)
            out.append(g.build(total) if isinstance(spec, schedules.CtrlSpec) else g)
        return out


class StronglyEntanglingLayerBuilt(BuiltAnsatz):
    """Built SEL (``:80-133``): its gates, built, in ``mods`` (a
    ``torch.nn.Sequential``); the ring offset runs modulo ``len(qubits) - 1``."""

    def __init__(self, num_qubits: int, qubits: typing.List[int], depth: int, rotations, q_params: torch.Tensor,
                 remapping: typing.Union[typing.Callable, None]):
        super().__init__()
        self.num_qubits, self.qubits, self.depth, self.rots = num_qubits, qubits, depth, rotations
        specs = schedules.strongly_entangling(qubits, depth, rotations, len(qubits) - 1, num_qubits, op.CNOT)
        self.mods = torch.nn.Sequential(*[_unbuilt(s, q_params, remapping).build(num_qubits) for s in specs])

    def forward(self, state: torch.Tensor):
        return self.mods(state)

    def __str__(self) -> str:
        return "SEL"

    def decompose(self) -> typing.List[op.Operator]:
        return [copy.deepcopy(m) for m in self.mods]
