"""MeasureProbability (``src/qandle/measurements.py:94-126``).

P(qubit q = 0) per sample as float32 (B, n).  The reference repeats the state
n times, masks and reduces (``:120-123``, an n x state temporary); the engine
runs one deterministic |psi|^2 reduction (``qs_measure_prob``), and inside a
Circuit fuses it into the last pass.  The reference's squeeze quirks are kept:
an unbatched state gives (n,), and so does a batched state with B = 1.
"""
from __future__ import annotations

import random
import warnings

import torch

from . import operators as op
from . import qasm
from .errors import UnbuiltGateError
from ._engine import plain

__all__ = ["MeasureProbability", "MeasureProbabilityBuilt", "BuiltMeasurement", "UnbuiltMeasurement",
           "MeasureAllAbsolute", "MeasureJointProbability", "MeasureState", "MeasurementDummy"]


def reference_shape(probs_bn: torch.Tensor) -> torch.Tensor:
    """Apply ``.squeeze().transpose(-1, 0)`` to the (n, B) layout, as the reference does."""
    return probs_bn.transpose(0, 1).squeeze().transpose(-1, 0)


class MeasurementDummy(op.UnbuiltOperator, torch.nn.Module):
    """One qubit of a decomposed measurement (``src/qandle/measurements.py:30-54``):
    a no-op on the state that keeps its place in drawings and OpenQASM."""

    def __init__(self, qubit: int):
        torch.nn.Module.__init__(self)
        self.qubit = qubit

    def __str__(self) -> str:
        return "M"

    def to_qasm(self) -> qasm.QasmRepresentation:
        return qasm.QasmRepresentation(gate_str=f"measure q[{self.qubit}] -> c[{self.qubit}]",
                                       qasm3_outputs=f"measured_{hash(random.random())}_{self.qubit}")

    def forward(self, state: torch.Tensor) -> torch.Tensor:
        warnings.warn("Forward on a decomposed measurement is a no-op. "
                      "Consider acting on the resulting statevector directly")
        return state

    def build(self, num_qubits: int, **kwargs) -> "MeasurementDummy":
        return self


class BuiltMeasurement(op.BuiltOperator):
    def __str__(self) -> str:
        return "M"

    def to_qasm(self):
        """``measure q[w] -> c[w];`` for every qubit, each an OpenQASM 3 output
        (``src/qandle/measurements.py:16-22``)."""
        i = hash(random.random())
        return [qasm.QasmRepresentation(gate_str=f"measure q[{w}] -> c[{w}]", qasm3_outputs=f"measured_{i}_{w}")
                for w in range(self.num_qubits)]

    def decompose(self):
        """One no-op ``MeasurementDummy`` per qubit (``src/qandle/measurements.py:24-54``): a
        decomposed circuit returns its state, as the reference's does."""
        warnings.warn("Decomposed measurements a no-op. Consider acting on the resulting statevector directly",
                      RuntimeWarning)
        return [MeasurementDummy(w) for w in range(self.num_qubits)]

    def to_matrix(self, **kwargs):
        return torch.eye(2 ** self.num_qubits)


class UnbuiltMeasurement(op.UnbuiltOperator):
    def __str__(self) -> str:
        return "M"

    def to_qasm(self):
        raise UnbuiltGateError(f"Unbuilt {self.__class__.__name__} cannot be converted to qasm.")

    def build(self, num_qubits: int, **kwargs) -> BuiltMeasurement:
        return BUILT_CLASS_RELATION[self.__class__](num_qubits=num_qubits, **kwargs)


class MeasureAllAbsolute(BuiltMeasurement):
    """|amplitude|^2 of every basis state (``src/qandle/measurements.py:73-82``),
    float32 with the state's shape (kernel ``qs_abs2``)."""

    def forward(self, state: torch.Tensor) -> torch.Tensor:
        from ._engine.lib import require_device

        require_device(state, "state")
        return plain.Abs2Fn.apply(state)


MeasureJointProbability = MeasureAllAbsolute  # alias, as in the reference


class MeasureState(BuiltMeasurement):
    """The state itself (``src/qandle/measurements.py:85-91``): a no-op."""

    def forward(self, state: torch.Tensor) -> torch.Tensor:
        return state


class MeasureProbability(UnbuiltMeasurement):
    r"""Probability of each qubit being in :math:`|0\rangle`."""


class MeasureProbabilityBuilt(BuiltMeasurement):
    def __init__(self, num_qubits: int):
        super().__init__()
        self.num_qubits = num_qubits
        self.qubits = list(range(num_qubits))

    def forward(self, state: torch.Tensor):
        from ._engine.lib import require_device

        require_device(state, "state")
        n = self.num_qubits
        if state.shape[-1] != 2 ** n:
            raise RuntimeError(f"state has {state.shape[-1]} amplitudes, expected {2 ** n}")
        st = state.reshape(-1, 2 ** n)
        probs = plain.MeasureProbFn.apply(st, n)
        return reference_shape(probs)


BUILT_CLASS_RELATION = {MeasureProbability: MeasureProbabilityBuilt}
