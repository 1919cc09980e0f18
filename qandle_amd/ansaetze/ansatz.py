"""Ansatz base classes (``src/qandle/ansaetze/ansatz.py:9-33``)."""
from __future__ import annotations

import abc
import typing

from .. import operators as op
from .. import utils


class Ansatz(op.Operator, abc.ABC):
    @abc.abstractmethod
    def decompose(self) -> typing.List[op.Operator]:
        """The gates the ansatz is made of."""

    def to_qasm(self):
        """The statements of its gates in order (``src/qandle/ansaetze/specialunitary.py:38-39, 113-114``,
        ``twolocal.py:43-44, 108-114``, ``stronglyentangling.py:126-127``,
        ``stronglyentangling_budget.py:83-84``)."""
        return [g.to_qasm() for g in self.decompose()]


class UnbuiltAnsatz(Ansatz, op.UnbuiltOperator, abc.ABC):
    """An ansatz exposed to the user; built into a :class:`BuiltAnsatz`."""

    @abc.abstractmethod
    def build(self, num_qubits: int, **kwargs) -> "BuiltAnsatz":
        """Builds the ansatz for a circuit of ``num_qubits``."""


class BuiltAnsatz(Ansatz, op.BuiltOperator, abc.ABC):
    """A built ansatz.  ``to_matrix`` multiplies out its decomposition
    (``src/qandle/ansaetze/ansatz.py:25-33``; a dense debugging aid)."""

    def to_matrix(self, **kwargs):
        gates = [g.build(num_qubits=self.num_qubits) if hasattr(g, "build") else g for g in self.decompose()]
        return utils.reduce_dot(*[g.to_matrix(**kwargs) for g in gates])
