"""CNOTs that end a measured circuit, folded into its readout.

``fused.readout_fold`` replaces the trailing CNOT network of a measured
program by one parity mask per measured bit.  The masks are checked against
the network itself on every basis state (brute force, small widths), and
through the fp64 oracle: the parity readout of the state before the CNOTs
equals the oracle's MeasureProbability after them.  The generated pass and
adjoint kernels that read the masks compile (host hipRTC); their numerical
parity runs on the GPU (test_gpu_specialised.py).
"""
import numpy as np
import pytest

from oracle import fp64
from qandle_amd._engine import fused, jit
from qandle_amd._engine import planner as PL


def _random_network(rng, n, count):
    out = []
    for _ in range(count):
        c, t = rng.choice(n, 2, replace=False)
        out.append(("cnot", (int(c), int(t))))
    return out


def _parity(x, m):
    return bin(x & m).count("1") & 1


@pytest.mark.parametrize("n,count,seed", [(2, 3, 0), (3, 6, 1), (5, 12, 2), (7, 20, 3), (7, 1, 4)])
def test_masks_follow_the_network(n, count, seed):
    rng = np.random.default_rng(seed)
    net = _random_network(rng, n, count)
    prog = fused.program_from_spec(n, [("ry", 0, 0.4)] + net)
    m, masks = fused.readout_fold(prog)
    assert m == 1 and len(masks) == n
    for x in range(2 ** n):
        y = x
        for _, (c, t) in net:
            if (y >> (n - 1 - c)) & 1:
                y ^= 1 << (n - 1 - t)
        for p in range(n):
            assert (y >> p) & 1 == _parity(x, masks[p]), (x, p)


def test_ladder_masks_are_prefix_parities():
    # CNOT(q, q + 1) for q = 0 .. n - 2: measured qubit q is the parity of qubits 0..q
    n = 9
    prog = fused.program_from_spec(n, [("rx", 3, 1.0)] + [("cnot", (q, q + 1)) for q in range(n - 1)])
    _, masks = fused.readout_fold(prog)
    for q in range(n):
        expect = sum(1 << (n - 1 - k) for k in range(q + 1))
        assert masks[n - 1 - q] == expect


@pytest.mark.parametrize("n,seed", [(4, 0), (6, 1), (8, 2)])
def test_parity_readout_matches_oracle(n, seed):
    rng = np.random.default_rng(seed)
    B = 3
    X = rng.uniform(0, np.pi, (B, n))
    body = []
    for _ in range(3):
        for q in range(n):
            body.append((str(rng.choice(["rx", "ry", "rz"])), q, float(rng.uniform(0, 2 * np.pi))))
        body += [("cnot", (q, q + 1)) for q in range(n - 1)]
    body += _random_network(rng, n, 5)
    prog = fused.program_from_spec(n, body, ("x", list(range(n)), "x"))
    m, masks = fused.readout_fold(prog)
    assert all(g[0] != "cnot" for g in body[m - 1:m])
    ops = [("embed", list(range(n)), "x", X)] + [(g[0], g[1], g[2]) if g[0] != "cnot" else g for g in body]

    def run(ops_):
        psi = fp64.embed(X, n, list(range(n)), "x")
        for op in ops_[1:]:
            if op[0] == "cnot":
                psi = fp64.apply_cnot(psi, n, *op[1])
            else:
                psi = fp64.apply_1q(psi, n, op[1], fp64.rot_matrix(op[0][1], op[2]))
        return psi

    full = fp64.measure_prob(run(ops), n)
    pre = np.abs(run(ops[: 1 + m])) ** 2
    idx = np.arange(2 ** n)
    folded = np.empty((B, n))
    for q in range(n):
        par = np.array([_parity(int(x), masks[n - 1 - q]) for x in idx])
        folded[:, q] = pre[:, par == 0].sum(axis=1)
    np.testing.assert_allclose(folded, full, rtol=0, atol=1e-12)


def test_no_fold_cases():
    n = 4
    assert fused.readout_fold(fused.program_from_spec(n, [("cnot", (0, 1)), ("rx", 0, 0.1)])) is None
    assert fused.readout_fold(fused.program_from_spec(n, [("cnot", (0, 1)), ("cnot", (1, 2))])) is None
    assert fused.readout_fold(fused.program_from_spec(n, [("rx", 0, 0.1), ("cnot", (0, 1))], measure=False)) is None
    assert not fused.folds(fused.program_from_spec(n, [("rx", 0, 0.1), ("cnot", (0, 1))]), False)


def test_folded_plan_drops_the_trailing_cnots():
    n = 14
    body = [("ry", q, 0.3) for q in range(n)] + [("cnot", (q, q + 1)) for q in range(n - 1)]
    prog = fused.program_from_spec(n, body)
    full = fused.get_plan(prog, 5, 8, fused.DEFAULT_COST, False)
    folded = fused.get_plan(prog, 5, 8, fused.DEFAULT_COST, True)
    assert full is not folded
    assert sum(len(p.gates) for p in folded.passes) == n
    assert sum(len(p.gates) for p in full.passes) == 2 * n - 1


@pytest.mark.parametrize("n,R,TB", [(6, 5, 6), (12, 5, 6), (14, 5, 8), (9, 4, 8)])
def test_readout_kernels_compile(n, R, TB, tmp_path, monkeypatch):
    monkeypatch.setattr(jit, "CACHE_DIR", str(tmp_path))
    rng = np.random.default_rng(n)
    body = [("rx", q, 0.2 + q) for q in range(n)] + [("cnot", (q, q + 1)) for q in range(n - 1)]
    body += _random_network(rng, n, 4)
    prog = fused.program_from_spec(n, body, ("y", list(range(n)), "x"))
    _, masks = fused.readout_fold(prog)
    pl = fused.get_plan(prog, R, TB, PL.CostModel(budget=48.0, min_low=2), True)
    B = 3
    ws = fused.pass_words(prog, pl, B, "cpu")
    last = len(ws.passes) - 1
    srcs = [jit.pass_source(f"qs_pass_{i}", w.numpy(), n, R, TB, lpos, i == 0, i == last, readout=masks)
            for i, (w, _nsub, lpos) in enumerate(ws.passes)]
    assert "tot - " in srcs[last] and f"* {n + 1};" in srcs[last]
    aws, _ = fused.adjoint_words(prog, pl, B, "cpu", True)
    asrcs = [jit.adjoint_source(f"qs_adj_{i}", w.numpy(), n, R, TB, lpos, i == 0, readout=masks)
             for i, (w, _nsub, lpos) in enumerate(aws.passes)]
    assert "wz +=" in asrcs[0]
    assert len(jit._code_for("\n\n".join(srcs))) > 1000
    assert len(jit._code_for("\n\n".join(asrcs))) > 1000
