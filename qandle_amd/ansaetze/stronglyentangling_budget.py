"""StronglyEntanglingLayerBudget (``src/qandle/ansaetze/stronglyentangling_budget.py:16-90``).

A layer sized by its parameter count: the gate pattern is
:func:`schedules.budget_layers`.  Built at construction like the reference
(it is a built ansatz); each rotation draws its own ``torch.rand(1)`` angle
when it is built, in schedule order.
"""
from __future__ import annotations

import typing

import torch

from .. import config
from .. import operators as op
from .. import qw_map
from .. import utils
from . import schedules
from .ansatz import BuiltAnsatz

__all__ = ["StronglyEntanglingLayerBudget"]


class StronglyEntanglingLayerBudget(BuiltAnsatz):
    """Rotation layers cycling through ``rotations`` until ``param_budget``
    rotations are placed, with a ring of ``control_gate`` (CNOT or CZ) before
    every ``control_gate_spacing``-th layer.  The built gates are ``layers``."""

    def __init__(self, num_qubits_total: int, qubits: typing.Union[typing.List[int], None] = None,
                 param_budget: int = 1, control_gate=op.CNOT, control_gate_spacing=1, rotations=("rz", "ry"),
                 remapping: typing.Union[typing.Callable, None] = config.DEFAULT_MAPPING):
        super().__init__()
        if control_gate not in (op.CNOT, op.CZ):
            raise AssertionError(f"control gate {control_gate} is neither CNOT nor CZ")
        self.num_qubits = num_qubits_total
        self.qubits = list(range(num_qubits_total)) if qubits is None else qubits
        self.rots = [utils.parse_rot(r) for r in rotations]
        self.remapping = qw_map.none if remapping is None else remapping
        self.param_budget, self.control_gate, self.control_gate_spacing = param_budget, control_gate, control_gate_spacing
        specs = schedules.budget_layers(self.qubits, param_budget, self.rots, control_gate_spacing, control_gate)
        self._unbuilt = [
            s.gate(control=s.control, target=s.target) if isinstance(s, schedules.CtrlSpec)
            else s.rot(qubit=s.qubit, remapping=self.remapping)
            for s in specs
        ]
        self.layers = torch.nn.Sequential(*[g.build(num_qubits=num_qubits_total) for g in self._unbuilt])

    @property
    def layers_ub(self) -> list:
        """The unbuilt gates in order (the reference's attribute of that name)."""
        return self._unbuilt

    def __str__(self) -> str:
        return "SEL"

    def decompose(self) -> typing.List[op.UnbuiltOperator]:
        return self._unbuilt

    def forward(self, x):
        return self.layers(x)
