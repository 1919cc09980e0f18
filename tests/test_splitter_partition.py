"""The splitter's partition, pinned against walks of the reference algorithm
done by hand for the circuits of src/qandle/test/test_splitter.py:34-101.

Algorithm (src/qandle/splitter/main.py:28-55, grouping.py:30-64): one DAG node
per CNOT, an edge from each CNOT to the next CNOT on each of its qubits
(successors in the order the edges were added: qubit 0 first); while a node is
ungrouped, start a group at the first ungrouped node of least in-degree
(circuit order breaks ties) and grow it depth-first into successors whose
parents are all grouped while the group stays within max_qubits.  Each
one-qubit gate joins the group of the last CNOT before it on its qubit, or
the first one after it (main.py:58-105); each group is renumbered over its
sorted qubits (main.py:114-157).
"""
import pytest
import torch

import qandle_amd as qandle
from qandle_amd import config

op = qandle.operators


def _cnot10():
    pairs = [(0, 1), (0, 2), (1, 2), (2, 3), (3, 4), (4, 3), (3, 4), (4, 3), (3, 4), (4, 3), (3, 4), (4, 3),
             (3, 5), (5, 6), (6, 7), (6, 1), (7, 8), (8, 9)]
    return qandle.Circuit(num_qubits=10, layers=[op.CNOT(c, t) for c, t in pairs]), pairs


def _groups(split_circuit):
    """Per sub-circuit: (parent qubits in order, gates as (kind, parent qubits...))."""
    out = []
    for sc in split_circuit.circuit.subcircuits:
        inv = sc.inverse
        gates = []
        for m in sc.layers:
            if isinstance(m, op.BuiltCNOT):
                gates.append(("cnot", inv[m.c], inv[m.t]))
            else:
                gates.append((type(m).__name__.replace("Built", "").lower(), inv[m.qubit]))
        out.append((sorted(sc.mapping), gates))
        assert list(sc.mapping.values()) == list(range(len(sc.mapping)))
    return out


def test_cnot10_partition():
    circ, pairs = _cnot10()
    got = _groups(circ.split(max_qubits=5))
    # hand walk: group 0 grows from CNOT 0 through CNOTs 1..11 (qubits 0-4; CNOT 12 would add a
    # sixth qubit, CNOT 15 waits for CNOT 14); group 1 starts at CNOT 12 and takes 13, 14, 15
    # (qubits 1, 3, 5, 6, 7; CNOT 16 would bring qubit 8, a sixth); group 2 is CNOTs 16, 17
    expect = [
        ([0, 1, 2, 3, 4], [("cnot", c, t) for c, t in pairs[:12]]),
        ([1, 3, 5, 6, 7], [("cnot", c, t) for c, t in pairs[12:16]]),
        ([7, 8, 9], [("cnot", c, t) for c, t in pairs[16:]]),
    ]
    assert got == expect


def _su3_layers():
    su = qandle.SU(qubits=list(range(3)), reps=2, rotations=["rx", "ry"])
    return qandle.Circuit(layers=su.decompose()).decompose()


def test_small_su_partitions():
    """The unbuilt SU decomposes rotation-major: RX0 RX1 RX2 RY0 RY1 RY2, CNOT(0,1) CNOT(1,2),
    the six rotations, the two CNOTs, the six rotations (layers 0-21)."""
    c = _su3_layers()
    one = _groups(c.split(max_qubits=3))
    assert len(one) == 1 and one[0][0] == [0, 1, 2] and len(one[0][1]) == 22
    assert _groups(c.split(max_qubits=4)) == one
    # max 2: no CNOT can join another, four groups {6}, {7}, {14}, {15}; each rotation goes to the
    # last CNOT before it on its qubit, else the first after it
    got = _groups(c.split(max_qubits=2))
    rx, ry = "rx", "ry"
    expect = [
        ([0, 1], [(rx, 0), (rx, 1), (ry, 0), (ry, 1), ("cnot", 0, 1), (rx, 0), (ry, 0)]),
        ([1, 2], [(rx, 2), (ry, 2), ("cnot", 1, 2), (rx, 1), (rx, 2), (ry, 1), (ry, 2)]),
        ([0, 1], [("cnot", 0, 1), (rx, 0), (ry, 0)]),
        ([1, 2], [("cnot", 1, 2), (rx, 1), (rx, 2), (ry, 1), (ry, 2)]),
    ]
    assert got == expect


def test_ten_qubit_su_partition():
    """Built SU(reps=2, rx/ry) on 10 qubits decomposes as block 0, block 1, chain a, block 2,
    chain b (src/qandle/ansaetze/specialunitary.py:116-123).  Hand walk with max 5:
    group 0 = a0..a3 + b0..b2 (qubits 0-4), group 1 = a4..a6 + b3..b5 (qubits 3-7),
    group 2 = a7, a8 + b6..b8 (qubits 6-9)."""
    su = qandle.SU(reps=2, rotations=["rx", "ry"], qubits=list(range(10)))
    c = qandle.Circuit(layers=[su]).decompose()
    got = _groups(c.split(max_qubits=5))
    assert [g[0] for g in got] == [[0, 1, 2, 3, 4], [3, 4, 5, 6, 7], [6, 7, 8, 9]]
    cn = [[(t[1], t[2]) for t in g[1] if t[0] == "cnot"] for g in got]
    assert cn == [[(0, 1), (1, 2), (2, 3), (3, 4), (0, 1), (1, 2), (2, 3)],
                  [(4, 5), (5, 6), (6, 7), (3, 4), (4, 5), (5, 6)],
                  [(7, 8), (8, 9), (6, 7), (7, 8), (8, 9)]]
    rot_qubits = [sorted({t[1] for t in g[1] if t[0] != "cnot"}) for g in got]
    # blocks 0 and 1 follow the first CNOT on their qubit, block 2 the last CNOT of chain a
    assert rot_qubits == [[0, 1, 2, 3, 4], [4, 5, 6, 7], [7, 8, 9]]
    assert [sum(t[0] != "cnot" for t in g[1]) for g in got] == [28, 18, 14]
    assert sum(len(g[1]) for g in got) == 78  # every gate lands in exactly one group


def test_max_qubits_guard():
    circ, _ = _cnot10()
    with pytest.raises(AssertionError):
        circ.split(max_qubits=1)


def test_split_remapping_rule():
    """Rotations in a sub-circuit keep their remapping (so split == unsplit); the
    reference's rebuild (main.py:143-153) resets it to DEFAULT_MAPPING, which
    SPLIT_KEEP_REMAPPING = False reproduces."""
    layers = [op.RX(0, theta=0.3, remapping=None), op.CNOT(0, 1), op.RY(1, theta=0.7, remapping=None),
              op.CNOT(1, 2), op.RZ(2, theta=1.1), op.CNOT(2, 0)]
    kept = qandle.Circuit(layers, num_qubits=3, split_max_qubits=2)
    rots = [m for sc in kept.circuit.subcircuits for m in sc.layers if hasattr(m, "remapping")]
    assert [m.remapping for m in rots] == [qandle.qw_map.none, qandle.qw_map.none, config.DEFAULT_MAPPING]
    old = config.SPLIT_KEEP_REMAPPING
    try:
        config.SPLIT_KEEP_REMAPPING = False
        reset = qandle.Circuit(layers, num_qubits=3, split_max_qubits=2)
    finally:
        config.SPLIT_KEEP_REMAPPING = old
    rots = [m for sc in reset.circuit.subcircuits for m in sc.layers if hasattr(m, "remapping")]
    assert all(m.remapping is config.DEFAULT_MAPPING for m in rots)
    # the thetas are the same Parameters either way (the rebuild passes theta through)
    assert [float(m.theta.detach()) for m in rots] == pytest.approx([0.3, 0.7, 1.1])


def test_custom_gate_in_a_split():
    m = torch.tensor([[0.6, 0.8j], [0.8j, 0.6]], dtype=torch.complex64)
    layers = [op.CNOT(0, 1), op.CustomGate(qubit=2, matrix=m, num_qubits=3, self_description="V"), op.CNOT(1, 2),
              op.CNOT(2, 0)]
    c = qandle.Circuit(layers, num_qubits=3, split_max_qubits=2)
    subs = c.circuit.subcircuits
    us = [(sc, g) for sc in subs for g in sc.layers if isinstance(g, op.CustomGate)]
    assert len(us) == 1
    sc, g = us[0]
    assert g.num_qubits == len(sc.mapping) and sc.inverse[g.qubit] == 2 and g.description == "V"
    assert torch.equal(torch.as_tensor(g.original_matrix), m)
