"""StronglyEntanglingLayerBudget (``src/qandle/ansaetze/stronglyentangling_budget.py:16-90``)."""
from __future__ import annotations

import typing

import torch

from .. import config
from .. import operators as op
from .. import qw_map
from .. import utils
from .ansatz import BuiltAnsatz

__all__ = ["StronglyEntanglingLayerBudget"]


class StronglyEntanglingLayerBudget(BuiltAnsatz):
    """Rotations on every qubit, cycling through ``rotations``, until
    ``param_budget`` parameters are placed, with a ring of control gates (CNOT
    or CZ) every ``control_gate_spacing`` layers (``:16-90``).  Built on
    construction; its gates are ``layers`` (a ``torch.nn.Sequential``)."""

    def __init__(self, num_qubits_total: int, qubits: typing.Union[typing.List[int], None] = None,
                 param_budget: int = 1, control_gate=op.CNOT, control_gate_spacing=1, rotations=("rz", "ry"),
                 remapping: typing.Union[typing.Callable, None] = config.DEFAULT_MAPPING):
        super().__init__()
        assert control_gate in [op.CNOT, op.CZ], f"Control gate {control_gate} not supported"
        if qubits is None:
            qubits = list(range(num_qubits_total))
        self.qubits = qubits
        self.num_qubits = num_qubits_total
        self.rots = [utils.parse_rot(r) for r in rotations]
        if remapping is None:
            remapping = qw_map.none
        self.remapping = remapping
        self.param_budget = param_budget
        self.control_gate = control_gate
        self.control_gate_spacing = control_gate_spacing
        self.layers_ub = self._allocate_params()
        self.layers = torch.nn.Sequential(*[layer.build(num_qubits=num_qubits_total) for layer in self.layers_ub])

    def _allocate_params(self):
        out = []
        current_params, current_depth, control_iteration = 0, 0, -1
        k = 0
        while True:
            r = self.rots[k % len(self.rots)]
            k += 1
            if current_depth % self.control_gate_spacing == 0 and current_depth > 0:
                control_iteration = (control_iteration + 1) % len(self.qubits)
                for ci in range(len(self.qubits)):
                    it = control_iteration % (len(self.qubits) - 1)
                    ti = (ci + 1 + it) % len(self.qubits)
                    out.append(self.control_gate(control=self.qubits[ci], target=self.qubits[ti]))
            current_depth += 1
            for q in self.qubits:
                current_params += 1
                if current_params > self.param_budget:
                    return out
                out.append(r(qubit=q, remapping=self.remapping))

    def __str__(self) -> str:
        return "SEL"

    def decompose(self) -> typing.List[op.UnbuiltOperator]:
        return self.layers_ub

    def forward(self, x):
        return self.layers(x)
