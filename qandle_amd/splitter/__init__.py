"""Circuit splitter (``src/qandle/splitter/main.py``, ``grouping.py``).

Host-side planning, restated without networkx: a CNOT dependency DAG
(``main.py:28-55``), greedy grouping into groups of at most ``max_qubits``
qubits (``grouping.py:30-64``), assignment of every single-qubit gate to the
group of the nearest CNOT on its qubit (``main.py:58-105``) and renumbering of
each group's qubits (``main.py:114-157``).  The split only changes how the
reference executes (smaller dense matrices); this engine lowers the
sub-circuits back onto parent bit positions, so a split circuit computes the
same result as the unsplit one.

One deliberate difference, switchable: a rotation's ``remapping`` is carried
into its sub-circuit (``config.SPLIT_KEEP_REMAPPING``, default True).  The
reference rebuilds every rotation from its class, theta and name only
(``main.py:143-153``), so the gate falls back to ``config.DEFAULT_MAPPING``
and a split circuit of ``remapping=None`` gates computes other angles than the
unsplit one; ``SPLIT_KEEP_REMAPPING = False`` reproduces that rebuild.
"""
from __future__ import annotations

import dataclasses
import typing

from .. import operators as op


@dataclasses.dataclass
class SubcircuitContainer:
    layers: list
    mapping: dict  # parent qubit -> sub-circuit qubit


def _is_cnot(layer) -> bool:
    return isinstance(layer, (op.CNOT, op.BuiltCNOT))


class _Node:
    __slots__ = ("c", "t", "idx", "group", "succ", "pred")

    def __init__(self, c, t, idx):
        self.c, self.t, self.idx = c, t, idx
        self.group = -1
        self.succ: list = []
        self.pred: list = []


def _graph(circuit):
    nodes = [_Node(la.c, la.t, i) for i, la in enumerate(circuit.layers) if _is_cnot(la)]
    assert len(nodes) > 0, "No CNOTs in the circuit, instead of splitting, just use separate circuits for each qubit"
    used = {n.c for n in nodes} | {n.t for n in nodes}
    assert len(used) == circuit.num_qubits, (
        "Circuit has qubits not used in CNOTs. Instead of splitting, just use separate circuits "
        f"for separate qubits. Only used qubits are: {used}"
    )
    per_qubit: typing.Dict[int, list] = {w: [] for w in range(circuit.num_qubits)}
    for n in nodes:
        per_qubit[n.c].append(n)
        per_qubit[n.t].append(n)
    for w in range(circuit.num_qubits):
        chain = per_qubit[w]
        for a, b in zip(chain, chain[1:]):
            if b not in a.succ:
                a.succ.append(b)
                b.pred.append(a)
    return nodes


def _group(nodes, max_qubits: int):
    assert max_qubits >= 2, "max_qubits must be >= 2"

    def grow(start, used: set):
        used.add(start.c)
        used.add(start.t)
        for nb in start.succ:
            if nb.group == -1 and all(p.group != -1 for p in nb.pred):
                if len(used | {nb.c, nb.t}) <= max_qubits:
                    nb.group = start.group
                    grow(nb, used)

    while any(n.group == -1 for n in nodes):
        first = next(n for n in sorted(nodes, key=lambda n: len(n.pred)) if n.group == -1)
        first.group = max(n.group for n in nodes) + 1
        grow(first, {first.c, first.t})


def _assign(nodes, layers):
    groups = {g: {} for g in sorted({n.group for n in nodes})}
    by_index = {n.idx: n for n in nodes}
    for i, la in enumerate(layers):
        if _is_cnot(la):
            groups[by_index[i].group][i] = la
    for i, la in enumerate(layers):
        if _is_cnot(la):
            continue
        w = la.qubit
        prev = nxt = None
        for j, other in enumerate(layers):
            if _is_cnot(other) and (other.c == w or other.t == w):
                if j < i:
                    prev = j
                elif j > i:
                    nxt = j
                    break
        near = prev if prev is not None else nxt
        assert near is not None, f"no CNOT touches qubit {w}"
        groups[by_index[near].group][i] = la
    return {g: [d[i] for i in sorted(d)] for g, d in groups.items()}


def _normalize(gates) -> SubcircuitContainer:
    qubits = set()
    for g in gates:
        if _is_cnot(g):
            qubits |= {g.c, g.t}
        else:
            qubits.add(g.qubit)
    mapping = {q: i for i, q in enumerate(sorted(qubits))}
    from .. import config

    keep_remap = getattr(config, "SPLIT_KEEP_REMAPPING", True)
    new_layers = []
    for g in gates:
        if _is_cnot(g):
            new_layers.append(op.CNOT(control=mapping[g.c], target=mapping[g.t]))
            continue
        if isinstance(g, op.CustomGate):
            # a custom gate is built already: rebuilt for the sub-register (main.py:131-138)
            new_layers.append(op.CustomGate(qubit=mapping[g.qubit], matrix=g.original_matrix,
                                            num_qubits=len(mapping), self_description=g.description))
            continue
        cls = op.BUILT_CLASS_RELATION.T[type(g)] if isinstance(g, op.BuiltOperator) else type(g)
        kwargs = {}
        if getattr(g, "theta", None) is not None:
            kwargs["theta"] = g.theta
        if getattr(g, "name", None) is not None:
            kwargs["name"] = g.name
        if keep_remap and hasattr(g, "remapping"):
            kwargs["remapping"] = g.remapping
        new_layers.append(cls(qubit=mapping[g.qubit], **kwargs))
    return SubcircuitContainer(new_layers, mapping)


def split(circuit, max_qubits: int = 5) -> typing.Dict[int, SubcircuitContainer]:
    """``src/qandle/splitter/main.py:12-25``."""
    assert circuit.num_qubits > 1, "Can't split circuits with only 1 qubit"
    circuit = circuit.decompose()
    assert all(isinstance(la, op.Operator) for la in circuit.layers), "Unknown layer type in circuit"
    nodes = _graph(circuit)
    _group(nodes, max_qubits)
    assigned = _assign(nodes, list(circuit.layers))
    return {g: _normalize(gates) for g, gates in assigned.items()}


__all__ = ["split", "SubcircuitContainer"]
