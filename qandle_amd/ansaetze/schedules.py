"""Gate schedules of the entangling ansätze, as data.

An ansatz here is a fixed pattern of one-qubit rotations and two-qubit
control gates.  The functions below only list that pattern, as
:class:`RotSpec` / :class:`CtrlSpec` records in application order; the ansatz
classes turn the records into operator modules (``build``), into unbuilt
operators (``decompose``), and the fused lowering then walks the built
modules.  Keeping the pattern in one place means the three views cannot
drift apart.

Patterns (the behaviour of the reference's classes, restated):

* StronglyEntanglingLayer, layer ``d`` (``src/qandle/ansaetze/stronglyentangling.py:60-75, 95-120``):
  every rotation of ``rotations`` on each qubit of ``qubits`` in order, the
  angle taken from ``q_params[d, wire index, rotation index]``; then a ring of
  CNOTs from ``qubits[i]`` to ``qubits[(i + 1 + shift_d) mod len(qubits)]``.
* StronglyEntanglingLayerBudget (``src/qandle/ansaetze/stronglyentangling_budget.py:50-75``):
  rotation layers that cycle through ``rotations`` and stop once
  ``param_budget`` rotations are placed; before every layer ``L >= 1`` with
  ``L`` a multiple of ``control_gate_spacing`` a control ring whose offset
  advances by one per ring.  The layer in which the budget runs out is still
  opened, so a budget that fills whole layers can end with a ring.
"""
from __future__ import annotations

import dataclasses
import typing


@dataclasses.dataclass(frozen=True)
class RotSpec:
    """One rotation: ``rot`` (an unbuilt rotation class) on ``qubit``; ``param``
    indexes the ansatz's parameter tensor, or is None when the gate draws its own."""

    rot: typing.Any
    qubit: int
    param: typing.Optional[typing.Tuple[int, ...]] = None


@dataclasses.dataclass(frozen=True)
class CtrlSpec:
    """One two-qubit control gate (an unbuilt class such as CNOT or CZ)."""

    gate: typing.Any
    control: int
    target: int


def ring(qubits: typing.Sequence[int], offset: int) -> typing.List[typing.Tuple[int, int]]:
    """(control, target) pairs of a ring over ``qubits``: wire i drives wire
    i + 1 + offset, indices modulo the ring size."""
    q = len(qubits)
    return [(qubits[i], qubits[(i + 1 + offset) % q]) for i in range(q)]


def sel_ring_offset(layer: int, period: int, num_qubits_total) -> int:
    """Offset of layer ``layer``'s CNOT ring: ``layer mod period``.  The
    reference checks that the offset leaves room in the register
    (``offset + 1 < num_qubits_total``) and fails with an AssertionError
    otherwise (``src/qandle/ansaetze/stronglyentangling.py:112``)."""
    offset = layer % period
    if not offset + 1 < num_qubits_total:
        raise AssertionError(f"CNOT ring offset {offset} does not fit a {num_qubits_total}-qubit register")
    return offset


def strongly_entangling(qubits, depth: int, rotations, period: int, num_qubits_total, cnot):
    """Records of a StronglyEntanglingLayer of ``depth`` layers.  ``period`` is the
    modulus of the ring offset: ``len(qubits) - 1`` for the built layer and the
    register size minus one for the unbuilt layer's decomposition, as in the
    reference (``src/qandle/ansaetze/stronglyentangling.py:75, 104``)."""
    out: typing.List[typing.Union[RotSpec, CtrlSpec]] = []
    for d in range(depth):
        out += [RotSpec(rot, w, (d, wi, ri)) for wi, w in enumerate(qubits) for ri, rot in enumerate(rotations)]
        offset = sel_ring_offset(d, period, num_qubits_total)
        out += [CtrlSpec(cnot, c, t) for c, t in ring(qubits, offset)]
    return out


def budget_layers(qubits, param_budget: int, rotations, spacing: int, ctrl):
    """Records of a StronglyEntanglingLayerBudget.

    With ``Q = len(qubits)`` the budget fills ``param_budget // Q`` whole
    rotation layers and ``param_budget % Q`` rotations of one more; layer ``L``
    uses ``rotations[L mod len(rotations)]``.  Every layer index ``L`` in
    ``1 .. param_budget // Q`` that ``spacing`` divides is preceded by ring
    number k (0, 1, 2, ...), whose offset is ``(k mod Q) mod (Q - 1)``."""
    q = len(qubits)
    full, rest = divmod(max(param_budget, 0), q)
    out: typing.List[typing.Union[RotSpec, CtrlSpec]] = []
    rings = 0
    for layer in range(full + 1):
        if layer % spacing == 0 and layer > 0:  # a spacing of 0 fails here, at the first layer
            offset = (rings % q) % (q - 1)
            out += [CtrlSpec(ctrl, c, t) for c, t in ring(qubits, offset)]
            rings += 1
        rot = rotations[layer % len(rotations)]
        out += [RotSpec(rot, w) for w in qubits[: q if layer < full else rest]]
    return out
