"""Fusion planner + word encoder, checked on the CPU through the kernel emulator.

The emulator executes the exact int32 words the GPU kernel receives; the
result must equal the FP64 oracle for the same circuit.  Covers tiles that hold
several samples (n < K), tiles inside one sample (n >= K, several tiles per
sample), CNOT controls outside the registers, RZ on non-register bits, and
dependency order across passes.
"""
import numpy as np
import pytest

from oracle import fp64
from qandle_amd._engine import planner as PL
from tests.fused_emulator import emulate_adjoint, emulate_program

KIND = {"rx": PL.RX, "ry": PL.RY, "rz": PL.RZ}


def lower(n, ops, B):
    """oracle ops -> (gates, cs table, efac)."""
    gates, cs = [], []
    efac = np.zeros((B, n, 2), dtype=np.complex128)
    efac[:, :, 0] = 1.0
    for op in ops:
        if op[0] == "embed":
            _, rot, qubits, X = op
            X = np.broadcast_to(np.asarray(X, dtype=np.float64), (B, len(qubits)))
            for j, q in enumerate(qubits):
                U = fp64.rot_matrix(rot, X[:, j])
                efac[:, n - 1 - q, 0] = U[:, 0, 0]
                efac[:, n - 1 - q, 1] = U[:, 1, 0]
        elif op[0] == "cnot":
            gates.append(PL.Gate(PL.CNOT, n - 1 - op[2], n - 1 - op[1]))
        else:
            gates.append(PL.Gate(KIND[op[0]], n - 1 - op[1], -1, len(cs)))
            t = op[2] / 2
            cs.append((np.cos(t), np.sin(t)))
    cs = np.asarray(cs if cs else [(1.0, 0.0)], dtype=np.float64)
    return gates, cs, efac


def random_ops(rng, n, depth, B, cnot_mode="ladder"):
    ops = [("embed", "x", list(range(n)), rng.uniform(0, np.pi, size=(B, n)))]
    for _ in range(depth):
        for q in range(n):
            for k in ("rx", "ry", "rz"):
                if rng.random() < 0.7:
                    ops.append((k, q, float(rng.uniform(0, 2 * np.pi))))
        if cnot_mode == "ladder":
            for q in range(n - 1):
                ops.append(("cnot", q, q + 1))
        else:
            for _ in range(n):
                c, t = rng.choice(n, 2, replace=False)
                ops.append(("cnot", int(c), int(t)))
    return ops


@pytest.mark.parametrize(
    "n,B,R,TB,mode",
    [
        (6, 5, 3, 4, "ladder"),    # n < K: multi-sample tiles, partial last tile
        (7, 3, 5, 2, "ladder"),    # n = K
        (10, 2, 3, 4, "ladder"),   # 8 tiles per sample
        (10, 2, 3, 4, "random"),   # long-range CNOTs
        (11, 2, 2, 3, "random"),   # tiny registers: many sub-passes and passes
        (12, 1, 5, 4, "ladder"),
    ],
)
def test_emulated_plan_matches_oracle(n, B, R, TB, mode):
    rng = np.random.default_rng(n * 31 + R)
    ops = random_ops(rng, n, 3, B, mode)
    gates, cs, efac = lower(n, ops, B)
    cm = PL.CostModel(budget=40.0, min_low=2)
    state, probs, pl = emulate_program(gates, n, B, R, TB, cs, efac, cm=cm)
    ref = fp64.simulate(n, ops)
    np.testing.assert_allclose(state, ref, atol=1e-12)
    np.testing.assert_allclose(probs, fp64.measure_prob(ref, n), atol=1e-12)
    if n >= R + TB:
        assert len(pl.passes) >= 2


def test_emulated_input_state_and_no_gates():
    rng = np.random.default_rng(4)
    n, B = 9, 2
    psi0 = rng.normal(size=(B, 2 ** n)) + 1j * rng.normal(size=(B, 2 ** n))
    gates, cs, efac = lower(n, [], B)
    state, probs, pl = emulate_program(gates, n, B, 3, 4, cs, efac, src=psi0)
    np.testing.assert_allclose(state, psi0, atol=1e-15)
    np.testing.assert_allclose(probs, fp64.measure_prob(psi0, n), atol=1e-12)


def test_plan_respects_dependencies():
    """Every gate runs exactly once and no gate runs before an earlier one on a shared position."""
    rng = np.random.default_rng(8)
    n = 16
    ops = random_ops(rng, n, 4, 1, "random")
    gates, _, _ = lower(n, ops, 1)
    pl = PL.plan(n, gates, 5, 8)
    order = [g for p in pl.passes for sp in p.subs for g in sp.gates]
    assert sorted(order) == list(range(len(gates)))
    when = {g: i for i, g in enumerate(order)}
    for i, gi in enumerate(gates):
        for j in range(i):
            if set(gates[j].positions) & set(gi.positions):
                assert when[j] < when[i]
    for p in pl.passes:
        assert len(p.L) == 13 and p.L == sorted(p.L) and p.L[0] == 0
        for sp in p.subs:
            assert set(sp.P) <= set(p.L)
            for g in sp.gates:
                assert set(gates[g].need) <= set(sp.P)


def test_pass_count_c2_family():
    """20 qubits, 10 layers of RX,RY,RZ + ladder: far fewer passes than gates."""
    n = 20
    gates = []
    slot = 0
    for _ in range(10):
        for q in range(n):
            for k in (PL.RX, PL.RY, PL.RZ):
                gates.append(PL.Gate(k, n - 1 - q, -1, slot))
                slot += 1
        for q in range(n - 1):
            gates.append(PL.Gate(PL.CNOT, n - 2 - q, n - 1 - q))
    pl = PL.plan(n, gates, 5, 8)
    stats = PL.plan_stats(pl, gates)
    print(f"C2 family: {len(gates)} gates in {len(pl.passes)} passes;", [s["gates"] for s in stats])
    assert len(pl.passes) <= 60


@pytest.mark.parametrize("n,B,R,TB,mode", [(6, 3, 3, 4, "ladder"), (10, 2, 3, 4, "random"), (9, 2, 2, 3, "ladder")])
def test_emulated_adjoint_matches_fp64(n, B, R, TB, mode):
    """v2 adjoint words + composed gradient generators reproduce the FP64 adjoint."""
    rng = np.random.default_rng(n * 7 + B)
    ops = random_ops(rng, n, 2, B, mode)
    gates, cs, efac = lower(n, ops, B)
    cm = PL.CostModel(budget=40.0, min_low=2)
    state, probs, pl = emulate_program(gates, n, B, R, TB, cs, efac, cm=cm)
    w = rng.normal(size=(B, n))
    grads, psi0, lam0 = emulate_adjoint(gates, n, B, R, TB, cs, state, w, pl,
                                        embed_bits=[n - 1 - q for q in range(n)], embed_rot=0)
    ref = fp64.adjoint_grads(n, ops, w)
    rot_ops = [op for op in ops[1:] if op[0] != "cnot"]
    rot_ref = [g for g, op in zip(ref[1:], ops[1:]) if op[0] != "cnot"]
    for r, (op, gr) in enumerate(zip(rot_ops, rot_ref)):
        got = grads[r].sum()
        np.testing.assert_allclose(got, gr, atol=1e-10)
    nrot = len(rot_ops)
    gx = np.stack([grads[nrot + j] for j in range(n)], 1)
    np.testing.assert_allclose(gx, ref[0], atol=1e-10)
    np.testing.assert_allclose(psi0.reshape(B, -1), fp64.simulate(n, ops[:1]), atol=1e-10)


def test_segmenter_fuses_runs_and_splits_at_cnots():
    gates = [PL.Gate(PL.RX, 0, -1, 0), PL.Gate(PL.RY, 0, -1, 1), PL.Gate(PL.CNOT, 1, 0),
             PL.Gate(PL.RZ, 1, -1, 2), PL.Gate(PL.RX, 0, -1, 3), PL.Gate(PL.RZ, 7, -1, 4)]
    phases, segs = PL.segment_subpass(PL.SubPass([0, 1], list(range(6))), gates)
    assert phases == [5]
    assert segs[0].bits == {0: [0, 1]} and segs[0].cnots == [2]
    assert segs[1].bits == {1: [3], 0: [4]} and segs[1].cnots == []


def _rot2(kind, c, s):
    if kind == PL.RX:
        return np.array([[c, -1j * s], [-1j * s, c]])
    if kind == PL.RY:
        return np.array([[c, -s], [s, c]])
    return np.array([[c - 1j * s, 0], [0, c + 1j * s]])


@pytest.mark.parametrize("mode", ["ladder", "random"])
def test_leading_rotations_go_into_the_product_factors(mode):
    """Rotations ahead of every CNOT on their qubit (control or target) act on
    an unentangled qubit: applied to its factor, the remaining program gives
    the oracle's state; a rotation after a CNOT that only controls its qubit
    stays in the program."""
    from qandle_amd._engine import fused

    rng = np.random.default_rng(11 if mode == "ladder" else 12)
    n, B = 8, 3
    ops = random_ops(rng, n, 3, B, mode)
    gates, cs, efac = lower(n, ops, B)
    prog = fused.Program(n, gates, [fused.Slot("value")] * len(cs), ("embed", 0, list(range(n)), "x"), True)
    rest, per_pos = fused.leading_rotations(prog)
    assert sum(len(v) for v in per_pos.values()) + len(rest.gates) == len(gates)
    touched = set()
    for g in gates:
        if g.kind == PL.CNOT:
            touched.update(g.positions)
        elif g.pos not in touched:
            assert any(h is g for h in per_pos[g.pos])
    ef = efac.copy()
    for pos, gs in per_pos.items():
        for g in gs:
            U = _rot2(g.kind, *cs[g.slot])
            ef[:, pos] = ef[:, pos] @ U.T
    state, probs, _ = emulate_program(rest.gates, n, B, 3, 4, cs, ef, cm=PL.CostModel(budget=40.0, min_low=2))
    ref = fp64.simulate(n, ops)
    np.testing.assert_allclose(state, ref, atol=1e-12)
    # the control side counts: RX on a CNOT's control after that CNOT does not commute with it
    g2 = [PL.Gate(PL.CNOT, 0, 1), PL.Gate(PL.RX, 1, -1, 0), PL.Gate(PL.RY, 2, -1, 1)]
    p2 = fused.Program(3, g2, [fused.Slot("value")] * 2, ("input",), True)
    rest2, per2 = fused.leading_rotations(p2)
    assert [g.kind for g in rest2.gates] == [PL.CNOT, PL.RX] and list(per2) == [2]


@pytest.mark.parametrize("mode", ["ladder", "random"])
def test_lookahead_plan_matches_oracle_and_scores_no_worse(mode):
    """The lookahead schedule (each pass's local set chosen by the score of the
    whole remaining schedule) is a valid plan: emulated, it gives the oracle's
    state; and it never scores worse than the greedy ones it is compared with."""
    rng = np.random.default_rng(21 if mode == "ladder" else 22)
    n, B, R, TB = 11, 2, 3, 4
    ops = random_ops(rng, n, 4, B, mode)
    gates, cs, efac = lower(n, ops, B)
    greedy = PL.CostModel(budget=1024.0, min_low=2)
    look = PL.CostModel(budget=1024.0, min_low=2, lookahead=True)
    state, probs, pl = emulate_program(gates, n, B, R, TB, cs, efac, cm=look)
    ref = fp64.simulate(n, ops)
    np.testing.assert_allclose(state, ref, atol=1e-12)
    np.testing.assert_allclose(probs, fp64.measure_prob(ref, n), atol=1e-12)
    score = lambda p: 4 * len(p.passes) + sum(len(s.subs) for s in p.passes)  # noqa: E731
    assert score(PL.plan(n, gates, R, TB, look)) <= score(PL.plan(n, gates, R, TB, greedy))


def test_lookahead_tries_one_fewer_always_local_position():
    """The lookahead planner also plans with positions 0-2 always local (tile
    runs of 2^3 amplitudes) and keeps that plan when it scores strictly lower:
    the 24-qubit ladder then runs in 7 passes instead of 9, while the
    20-qubit one keeps its 6 passes over positions 0-3."""
    import qandle_amd as qandle
    from qandle_amd import workloads
    from qandle_amd._engine import fused

    def plan_of(n, B):
        w = workloads.rxryrz_ladder(n=n, B=B, depth=10, seed=1)
        prog = fused.lower_circuit(qandle.Circuit(w.layers, num_qubits=n).circuit)
        p = fused.leading_rotations(prog)[0]
        return fused.get_plan(p, 5, 8, None, fused.folds(p, True))

    p24, p20 = plan_of(24, 4), plan_of(20, 64)
    assert len(p24.passes) == 7 and any(3 not in q.L for q in p24.passes)
    assert all({0, 1, 2} <= set(q.L) for q in p24.passes)
    assert len(p20.passes) == 6 and all({0, 1, 2, 3} <= set(q.L) for q in p20.passes)
