"""The engine's own encoded programs, checked on the CPU.

``fused.pass_words`` and ``fused.adjoint_words`` produce exactly the int32
words and table recipes that ``qs_fused_pass`` / ``qs_adjoint_pass`` /
``qs_compose_tables`` receive on the GPU.  Here they are built for circuits of
every lowering path the GPU tests use (AngleEmbedding, shared angles, named
per-sample and 0-d inputs, tiles holding several samples, tiles inside one
sample), walked with the kernels' cursor order to assert that every index
stays inside its table, then executed by the emulator and compared with the
FP64 oracle (probabilities, angle gradients, embedding gradients).
"""
import numpy as np
import pytest
import torch

import qandle_amd as qandle
from oracle import fp64
from qandle_amd._engine import fused
from qandle_amd._engine import planner as PL
from tests import fused_emulator as EM

ROT = {"x": 0, "y": 1, "z": 2}


def case(name, seed):
    """(n, B, layers, oracle ops, call inputs as numpy).  Module angles are
    float32 parameters, so the oracle gets the same float32-rounded values."""
    rng = np.random.default_rng(seed)
    if name == "c1_ring":
        n, B = 8, 5
        X = rng.uniform(0, np.pi, (B, n))
        layers = [qandle.AngleEmbedding(name="x", qubits=list(range(n)), rotation="rx")]
        ops = [("embed", "x", list(range(n)), X)]
        for _ in range(3):
            for q in range(n):
                a = float(np.float32(rng.uniform(0, 2 * np.pi)))
                layers.append(qandle.RY(q, theta=torch.tensor(a), remapping=None))
                ops.append(("ry", q, a))
            for q in range(n):
                layers.append(qandle.CNOT(q, (q + 1) % n))
                ops.append(("cnot", q, (q + 1) % n))
        return n, B, layers, ops, {"x": X}
    if name in ("ladder13", "ladder4"):
        n, B = (13, 3) if name == "ladder13" else (4, 6)
        X = rng.uniform(0, np.pi, (B, n))
        layers = [qandle.AngleEmbedding(name="x", qubits=list(range(n)), rotation="ry")]
        ops = [("embed", "y", list(range(n)), X)]
        for _ in range(2):
            for q in range(n):
                for k, cls in (("rx", qandle.RX), ("ry", qandle.RY), ("rz", qandle.RZ)):
                    a = float(np.float32(rng.uniform(0, 2 * np.pi)))
                    layers.append(cls(q, theta=torch.tensor(a), remapping=None))
                    ops.append((k, q, a))
            for q in range(n - 1):
                layers.append(qandle.CNOT(q, q + 1))
                ops.append(("cnot", q, q + 1))
        return n, B, layers, ops, {"x": X}
    if name == "named12":
        n, B = 12, 4
        X = rng.uniform(0, np.pi, (B, n))
        r = rng.uniform(0, 3, B)
        s0 = float(rng.uniform(0, 3))
        layers = [qandle.AngleEmbedding(name="x", qubits=list(range(n)), rotation="rz")]
        ops = [("embed", "z", list(range(n)), X)]
        for q in range(n):
            a = float(np.float32(rng.uniform(0, 6)))
            layers.append(qandle.RY(q, theta=torch.tensor(a), remapping=None))
            ops.append(("ry", q, a))
        layers += [qandle.RX(3, name="r", remapping=None), qandle.CNOT(3, 9), qandle.RZ(9, name="r"),
                   qandle.RY(0, name="s"), qandle.CNOT(11, 0), qandle.RZ(5, name="s"), qandle.RX(5, name="r")]
        ops += [("rx", 3, r), ("cnot", 3, 9), ("rz", 9, r), ("ry", 0, s0), ("cnot", 11, 0), ("rz", 5, s0),
                ("rx", 5, r)]
        return n, B, layers, ops, {"x": X, "r": r, "s": np.float64(s0)}
    raise KeyError(name)


def coefficients(prog, B, kwargs):
    index, per_sample, total, _ = fused.coefficient_layout(prog, B)
    cs = np.zeros((max(total, 1), 2))
    for i in prog.param_slots:
        sl = prog.slots[i]
        a = float(sl.module.theta.detach()) if sl.kind == "param" else sl.value
        cs[index[i]] = (np.cos(a / 2), np.sin(a / 2))
    for i in prog.named_slots:
        v = np.broadcast_to(np.asarray(kwargs[prog.slots[i].name], dtype=np.float64), (B,))
        cs[index[i]: index[i] + B, 0] = np.cos(v / 2)
        cs[index[i]: index[i] + B, 1] = np.sin(v / 2)
    return cs


def factors(prog, B, n, X):
    efac = np.zeros((B, n, 2), complex)
    efac[:, :, 0] = 1.0
    _, rot, bits, _ = prog.init
    for j, pos in enumerate(bits):
        U = fp64.rot_matrix("xyz"[rot], X[:, j])
        efac[:, pos, 0] = U[:, 0, 0]
        efac[:, pos, 1] = U[:, 1, 0]
    return efac


@pytest.mark.parametrize("name", ["c1_ring", "ladder13", "ladder4", "named12"])
def test_engine_words_in_bounds_and_match_fp64(name):
    n, B, layers, ops, kw = case(name, 7)
    circ = qandle.Circuit(layers + [qandle.MeasureProbability()], num_qubits=n)
    prog = fused.lower_circuit(circ.circuit)
    assert prog is not None and prog.init[0] == "embed"
    R, TB = fused.choose_shape(n)
    cm = fused.DEFAULT_COST
    pl = fused.get_plan(prog, R, TB, cm)
    fw = fused.pass_words(prog, pl, B, "cpu")
    adj, nmain = fused.adjoint_words(prog, pl, B, "cpu", True, cm)
    assert nmain == len(pl.passes)
    cs = coefficients(prog, B, kw)
    rows = fused.rotation_rows(prog)
    nrows = len(rows) + len(prog.init[2])

    # every index the kernels derive stays in bounds
    for ws, adjoint in ((fw, False), (adj, True)):
        rec, off = ws.recipes.numpy(), ws.offsets.numpy()[: ws.nrec]
        EM.check_recipes(rec, off, len(cs), ws.nu, ws.nq, adjoint)
        for w, nsub, _ in ws.passes:
            EM.check_words(w.numpy(), R, TB, n, B, len(cs), ws.nu, ws.nq, nrows if adjoint else 0, adjoint, nsub)

    # executing them gives the oracle's results
    efac = factors(prog, B, n, kw["x"])
    utab, _ = EM.compose(fw.recipes.numpy(), B, cs, fw.nu, fw.nq)
    state, probs = EM.run_forward_words([(w.numpy(), lp) for w, _, lp in fw.passes], B, n, R, TB, cs, utab, efac)
    ref = fp64.simulate(n, ops)
    np.testing.assert_allclose(state, ref, atol=1e-10)
    np.testing.assert_allclose(probs, fp64.measure_prob(ref, n), atol=1e-10)

    wt = np.random.default_rng(3).normal(size=(B, n))
    utab_a, qtab_a = EM.compose(adj.recipes.numpy(), B, cs, adj.nu, adj.nq)
    grads, _, _ = EM.run_adjoint_words([(w.numpy(), lp) for w, _, lp in adj.passes], B, n, R, TB, cs,
                                       utab_a, qtab_a, state, wt)
    gref = fp64.adjoint_grads(n, ops, wt)
    rot_ops = [(k, op) for k, op in enumerate(ops) if op[0] in ("rx", "ry", "rz")]
    assert len(rot_ops) == len(rows)
    for r, (k, op) in enumerate(rot_ops):
        got = grads.get(r, np.zeros(B))
        exp = gref[k]
        if np.ndim(exp) == 0:
            got = got.sum()
        np.testing.assert_allclose(got, exp, atol=1e-9, err_msg=f"row {r} ({op[0]} on {op[1]})")
    gx = np.stack([grads[len(rows) + j] for j in range(n)], 1)
    np.testing.assert_allclose(gx, gref[0], atol=1e-9)


@pytest.mark.parametrize("R,TB", [(4, 7), (4, 8), (5, 8)])
def test_uneven_batch_words_in_bounds_and_match_fp64(R, TB):
    """UnevenBatch encodes every circuit against one shared coefficient table and
    one shared pair of fused-unitary tables (absolute indices): each circuit's
    words, executed alone against those shared tables, reproduce its oracle
    (at each tile shape the item kernels take: tiles of several samples and
    tiles inside one sample)."""
    from qandle_amd import workloads
    from qandle_amd._engine.uneven import UnevenBatch

    wls = [workloads.random_uneven(1, n_lo=n, n_hi=n, d_lo=3, d_hi=3, seed=s, build_layers=False)[0]
           for n, s in ((5, 1), (12, 2), (9, 3), (14, 4))]
    ub = UnevenBatch.from_workloads(wls, "cpu", R=R, TB=TB)
    ub._build_backward()
    R, TB = ub.R, ub.TB
    phi = ub.phi.numpy().astype(np.float64)
    cs = np.stack([np.cos(phi / 2), np.sin(phi / 2)], 1)
    fwt, bwt = ub.fw_tables, ub.bw_tables
    EM.check_recipes(fwt.recipes.numpy(), fwt.offsets.numpy()[: fwt.nrec], len(cs), fwt.nu, fwt.nq, False)
    EM.check_recipes(bwt.recipes.numpy(), bwt.offsets.numpy()[: bwt.nrec], len(cs), bwt.nu, bwt.nq, True)
    utab, _ = EM.compose(fwt.recipes.numpy(), 1, cs, fwt.nu, fwt.nq)
    utab_a, qtab_a = EM.compose(bwt.recipes.numpy(), 1, cs, bwt.nu, bwt.nq)
    k = 0
    for ci, (w, c) in enumerate(zip(wls, ub.circs)):
        n, B = c.n, c.B
        passes = []
        for p in c.plan.passes:
            words = ub.fw_chunks[k]
            k += 1
            EM.check_words(words, R, TB, n, B, len(cs), fwt.nu, fwt.nq, 0, False, len(p.subs))
            passes.append((words, p.L))
        X = w.inputs["x"].numpy().astype(np.float64)
        efac = np.zeros((B, n, 2), complex)
        efac[:, :, 0] = 1.0
        _, rot, bits, _ = c.prog.init
        for j, pos in enumerate(bits):
            U = fp64.rot_matrix("xyz"[rot], X[:, j])
            efac[:, pos, 0], efac[:, pos, 1] = U[:, 0, 0], U[:, 1, 0]
        ops = [("embed", "xyz"[rot], list(w.embed[1]), X)]
        for g in w.gates:
            ops.append(("cnot", g[1][0], g[1][1]) if g[0] == "cnot" else (g[0], g[1], float(np.float32(g[2]))))
        state, probs = EM.run_forward_words(passes, B, n, R, TB, cs, utab, efac)
        ref = fp64.simulate(n, ops)
        np.testing.assert_allclose(state, ref, atol=1e-6, err_msg=f"circuit {ci}")
        # the adjoint steps of this circuit
        steps = ub.bw_steps[ci]
        start = sum(len(s) for s in ub.bw_steps[:ci])
        apasses = []
        for li, (p, _, nsub) in enumerate(steps):
            words = ub.bw_chunks[start + li]
            EM.check_words(words, R, TB, n, B, len(cs), bwt.nu, bwt.nq, c.rows, True, nsub)
            apasses.append((words, p.L))
        wt = np.random.default_rng(ci).normal(size=(B, n))
        grads, _, _ = EM.run_adjoint_words(apasses, B, n, R, TB, cs, utab_a, qtab_a, state, wt)
        gref = fp64.adjoint_grads(n, ops, wt)
        rot_k = [i for i, op in enumerate(ops) if op[0] in ("rx", "ry", "rz")]
        got = np.array([grads.get(r, np.zeros(B)).sum() for r in range(len(rot_k))])
        np.testing.assert_allclose(got, [gref[i] for i in rot_k], atol=1e-6, err_msg=f"circuit {ci}")
        gx = np.stack([grads[c.nrot + j] for j in range(c.d)], 1)
        np.testing.assert_allclose(gx, gref[0], atol=1e-6, err_msg=f"circuit {ci}")


def test_warm_builds_what_execute_reads(monkeypatch):
    """fused.warm encodes and uploads, for one program and batch, everything a
    forward reads from the host (the leading rotations' tables, the plan, the
    pass words and table recipes); warming again uploads nothing, and the
    words are kept on the host for the kernel compilers.  CPU tensors stand in
    for the device ones here."""
    from qandle_amd import workloads
    from qandle_amd._engine import lib as L

    w = workloads.rxryrz_ladder(n=9, B=3, depth=3, seed=4)
    prog = fused.lower_circuit(qandle.Circuit(w.layers, num_qubits=9).circuit)
    R, TB = fused.interpreted_shape(9)
    calls = []
    real = L.upload
    monkeypatch.setattr(L, "upload", lambda *a: calls.append(a) or real(*a))
    fused.warm(prog, 3, R, TB, fused.INTERPRETED_COST, False, "cpu")
    first = len(calls)
    assert first > 0
    fused.warm(prog, 3, R, TB, fused.INTERPRETED_COST, False, "cpu")
    assert len(calls) == first
    lead = fused.leading_rotations(prog)
    assert lead is not None and ("leading_tables", 3, "cpu") in prog.words_cache
    rest = lead[0]
    pl = fused.get_plan(rest, R, TB, fused.INTERPRETED_COST, False)
    ws = fused.pass_words(rest, pl, 3, "cpu")
    assert len(calls) == first
    assert [h[0].tolist() for h in ws.host_passes()] == [d[0].tolist() for d in ws.passes]
    up = L.upload([1, 2, 3], torch.int32, "cpu")
    assert up.device.type == "cpu" and up.dtype == torch.int32 and up.tolist() == [1, 2, 3]
