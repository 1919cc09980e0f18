"""Text circuit diagrams (``src/qandle/drawer.py:1-131``).

One text line per qubit, ``q0``, ``q1``, ..., with the gates appended left to
right; the characters and spacing are the reference's, so a diagram reads the
same in either package:

- a gate on one qubit (anything with a ``qubit`` attribute) writes its label
  on that wire and, unless ``config.DRAW_SHIFT_LEFT``, dashes of the same
  width on every other wire;
- CNOT and CZ write a control dot and a target mark (``⊕`` or a second dot)
  with a crossing (``┼``, ``config.DRAW_CROSS_BETWEEN_CNOT``) on the wires in
  between;
- any other layer writes its label on every wire;
- a split circuit draws one ``SUB<k>`` box per sub-circuit on the wires it
  spans and the sub-circuits' own diagrams side by side underneath.

Labels are ``str(gate)``; with ``config.DRAW_SHOW_VALUES`` off, digits, dots,
underscores, blanks and brackets are dropped from them.
"""
from __future__ import annotations

import re

from . import config

__all__ = ["draw"]

_PLACEHOLDER = "SUB{}"


def _label(gate) -> str:
    text = str(gate)
    return text if config.DRAW_SHOW_VALUES else re.sub(r"[\d._\s()]", "", text)


def _is_two_qubit_phase(gate):
    from . import operators as op

    return isinstance(gate, (op.CNOT, op.BuiltCNOT, op.CZ, op.BuiltCZ))


def _pad_to(wire: str, width: int) -> str:
    return wire + config.DRAW_DASH * (width - len(wire))


def _draw_two_qubit(wires, gate):
    from . import operators as op

    dash = config.DRAW_DASH
    c, t = gate.c, gate.t
    lo, hi = min(c, t), max(c, t)
    width = max(len(wires[w]) for w in range(lo, hi + 1))
    if config.DRAW_SHIFT_LEFT:
        wires[c] = _pad_to(wires[c], width)
        wires[t] = _pad_to(wires[t], width)
    wires[c] += dash + "⬤"
    wires[t] += dash + ("⊕" if isinstance(gate, (op.CNOT, op.BuiltCNOT)) else "⬤")
    for w in range(lo + 1, hi):
        if config.DRAW_SHIFT_LEFT:
            wires[w] = _pad_to(wires[w], width)
        wires[w] += dash + "┼" if config.DRAW_CROSS_BETWEEN_CNOT else dash * 2
    if not config.DRAW_SHIFT_LEFT:
        for w in range(len(wires)):
            if w < lo or w > hi:
                wires[w] += dash * 2


def draw(circuit) -> str:
    """Text diagram of an (unsplit or split) circuit's layers."""
    from . import qcircuit

    if isinstance(circuit, qcircuit.SplittedCircuit):
        return _draw_split(circuit)
    dash = config.DRAW_DASH
    n = circuit.num_qubits
    wires = [f"q{i}" for i in range(n)]
    for gate in circuit.layers:
        qubit = getattr(gate, "qubit", None)
        if isinstance(qubit, int):
            text = _label(gate)
            wires[qubit] += dash + text
            if not config.DRAW_SHIFT_LEFT:
                for w in range(n):
                    if w != qubit:
                        wires[w] += dash * (len(text) + 1)
        elif _is_two_qubit_phase(gate):
            _draw_two_qubit(wires, gate)
        else:
            text = _label(gate)
            for w in range(n):
                wires[w] += dash + text
    width = max(len(w) for w in wires) + 1
    return "\n".join(_pad_to(w, width) for w in wires)


def _draw_split(circuit) -> str:
    from . import qcircuit

    n = circuit.num_qubits
    wires = [f"q{i}" for i in range(n)]
    boxes = {}
    for k, sub in enumerate(circuit.subcircuits):
        tag = _PLACEHOLDER.format(k)
        spanned = set(sub.mapping.keys())
        for w in range(n):
            wires[w] += "-" + (tag if w in spanned else "-" * len(tag))
        inner = qcircuit.UnsplittedCircuit(num_qubits=len(spanned), layers=list(sub.layers))
        boxes[k] = tag + "\n" + draw(inner)
    body = "\n".join(w + "-" for w in wires)
    return body + "\nSubcircuits:\n" + _format_splitted(boxes, how="column")


def _format_splitted(sub_drawing: dict, how: str = "column") -> str:
    """The sub-circuit diagrams below each other (``row``) or side by side
    (``column``: each padded to its widest line plus ``config.DRAW_SPLITTED_PAD``)."""
    if how == "row":
        return "\n".join(sub_drawing.values())
    if how != "column":
        raise ValueError("how must be 'column' or 'row'.")
    blocks = [v.split("\n") for v in sub_drawing.values()]
    height = max(len(b) for b in blocks)
    blocks = [b + [""] * (height - len(b)) for b in blocks]
    widths = [max(len(line) for line in b) + config.DRAW_SPLITTED_PAD for b in blocks]
    rows = []
    for i in range(height):
        rows.append("".join(b[i].ljust(wd) for b, wd in zip(blocks, widths)) + "\n")
    return "".join(rows)
