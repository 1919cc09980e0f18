"""Run-time specialised pass kernels: the generated sources compile.

The generator turns a pass's encoded words into HIP source; here it is
compiled with hipRTC (host-only, no GPU needed) for the shapes the engine
uses: tiles of several samples and of part of one sample, the product-state
generator, the measuring last pass, register and uniform-control CNOTs, RZ
phases and per-sample unitaries.  Numerical parity of the kernels is tested
on the GPU (test_gpu_specialised.py).
"""
import os

import numpy as np
import pytest

import qandle_amd as qandle
from qandle_amd._engine import fused, jit
from qandle_amd._engine import planner as PL


def _program(n, named):
    layers = [qandle.AngleEmbedding(name="x", qubits=list(range(n)), rotation="rx")]
    for q in range(n):
        layers += [qandle.RX(q, theta=0.3 + q, remapping=None), qandle.RZ(q, theta=0.2, remapping=None)]
    for q in range(n - 1):
        layers.append(qandle.CNOT(q, q + 1))
    layers.append(qandle.CNOT(n - 1, 0))
    if named:
        layers += [qandle.RY(1, name="r", remapping=None), qandle.RZ(n - 1, name="r", remapping=None)]
    layers.append(qandle.MeasureProbability())
    circ = qandle.Circuit(layers, num_qubits=n)
    return fused.lower_circuit(circ.circuit)


@pytest.mark.parametrize("n,R,TB,named", [(6, 5, 6, True), (13, 4, 8, False), (12, 5, 6, True), (16, 3, 8, False)])
def test_generated_pass_kernels_compile(n, R, TB, named, tmp_path, monkeypatch):
    monkeypatch.setattr(jit, "CACHE_DIR", str(tmp_path))
    prog = _program(n, named)
    B = 3
    pl = fused.get_plan(prog, R, TB, PL.CostModel(budget=48.0, min_low=2))
    ws = fused.pass_words(prog, pl, B, "cpu")
    srcs, names = [], []
    for i, (w, _nsub, lpos) in enumerate(ws.passes):
        nm = f"qs_pass_{i}"
        srcs.append(jit.pass_source(nm, w.numpy(), n, R, TB, lpos, i == 0, i == len(ws.passes) - 1))
        names.append(nm)
    src = "\n\n".join(srcs)
    for nm in names:
        assert f"{nm}(JArgs A)" in src
    if R < n:
        assert "mfma_pair" in src
    code = jit._code_for(src)
    assert len(code) > 1000
    assert os.listdir(str(tmp_path))  # cached


@pytest.mark.parametrize("n,R,TB,named", [(6, 5, 6, True), (13, 4, 8, False), (12, 5, 6, True)])
def test_generated_adjoint_kernels_compile(n, R, TB, named, tmp_path, monkeypatch):
    monkeypatch.setattr(jit, "CACHE_DIR", str(tmp_path))
    prog = _program(n, named)
    B = 3
    pl = fused.get_plan(prog, R, TB, PL.CostModel(budget=48.0, min_low=2))
    ws, nmain = fused.adjoint_words(prog, pl, B, "cpu", True)
    srcs, names = [], []
    for i, (w, _nsub, lpos) in enumerate(ws.passes):
        nm = f"qs_adj_{i}"
        srcs.append(jit.adjoint_source(nm, w.numpy(), n, R, TB, lpos, i == 0))
        names.append(nm)
    src = "\n\n".join(srcs)
    assert "A.seed_w" in srcs[0] and "A.seed_w" not in srcs[-1]
    assert "im_trace" in src and "A.gpart" in src
    code = jit._code_for(src)
    assert len(code) > 1000


@pytest.mark.parametrize("opts", [{"lanes": "lds", "split_pairs": True}, {"lanes": True}, {"lanes": False,
                                                                                         "split_pairs": False},
                                  {"lanes": "lds", "split_pairs": True, "interleave": 4, "wave_sync": True,
                                   "stagger": 2}])
def test_pass_kernel_options_compile(opts, tmp_path, monkeypatch):
    """The operand-table and tile-move variants of the pass kernels (the engine
    default is lanes="lds" with split pairs) and the pipelined form compile."""
    monkeypatch.setattr(jit, "CACHE_DIR", str(tmp_path))
    n, R, TB = 14, 5, 8
    prog = _program(n, True)
    B = 3
    pl = fused.get_plan(prog, R, TB, PL.CostModel(budget=48.0, min_low=2))
    ws = fused.pass_words(prog, pl, B, "cpu")
    last = len(ws.passes) - 1
    off = 0
    srcs = []
    for i, (w, _nsub, lpos) in enumerate(ws.passes):
        srcs.append(jit.pass_source(f"qs_pass_{i}", w.numpy(), n, R, TB, lpos, i == 0, i == last, waves=2,
                                    lane_off=off, **opts))
        off += len(jit.lane_entries(w.numpy(), R, TB, False))
    if opts.get("lanes") == "lds":
        assert "ctab[" in srcs[0]
    srcs.append(jit.pass_source_pipelined("qs_pipe", ws.passes[1][0].numpy(), n, R, TB, ws.passes[1][2], False,
                                          False, 2, None, 8))
    assert len(jit._code_for("\n\n".join(srcs))) > 1000


def test_adjoint_lane_table_variant_compiles(tmp_path, monkeypatch):
    monkeypatch.setattr(jit, "CACHE_DIR", str(tmp_path))
    n, R, TB = 13, 4, 8
    prog = _program(n, False)
    B = 3
    pl = fused.get_plan(prog, R, TB, PL.CostModel(budget=48.0, min_low=2))
    ws, nmain = fused.adjoint_words(prog, pl, B, "cpu", True)
    off, srcs = 0, []
    for i, (w, _nsub, lpos) in enumerate(ws.passes):
        srcs.append(jit.adjoint_source(f"qs_adj_{i}", w.numpy(), n, R, TB, lpos, i == 0, waves=2, lanes="lds",
                                       split_pairs=True, lane_off=off))
        off += len(jit.adjoint_lane_entries(w.numpy(), R, TB))
    assert off > 0 and any("ctab[" in s_ for s_ in srcs)
    assert len(jit._code_for("\n\n".join(srcs))) > 1000


def test_lane_mode_keeps_the_compact_table_form():
    """Kernels built with lanes="lds" index a compact per-pass block, so the
    program must report "lds" (not merely a true value) for the caller to pack
    the matching table."""
    assert jit.lane_mode("lds", False) == "lds"
    assert jit.lane_mode(True, False) is True
    assert jit.lane_mode("lds", True) is False
    assert jit.lane_mode(False, False) is False


def test_lane_block_is_not_the_table_order_with_per_sample_unitaries():
    """With per-sample unitaries between the shared ones, a pass's compact lane
    block lists table entries out of the identity order: the case in which a
    whole-table operand array would hand the kernels wrong operands."""
    n, B = 12, 4
    prog = _program(n, True)
    R, TB = fused.choose_shape(n, grad=True)
    pl = fused.get_plan(prog, R, TB, None, fused.folds(prog, True))
    ws = fused.pass_words(prog, pl, B, "cpu")
    ents = [u for w, _ns, _lp in ws.passes for u in jit.lane_entries(w.numpy(), R, TB, n < R + TB)]
    assert ents and ents != list(range(len(ents)))


def test_hermitian_trace_matches_the_full_trace():
    """The adjoint kernels read Im Tr(Q M) for a generator Q = V P V^dagger
    (Hermitian) from four sums over amplitude pairs (herm_trace):
    Im M00, Im M11, Im M01 + Im M10 and Re M01 - Re M10, weighted by Q00,
    Q11, Re Q10 and Im Q10.  Here the same algebra in numpy against Im Tr(Q M)
    for random pairs, every Pauli generator and random unitaries V."""
    rng = np.random.default_rng(5)
    paulis = [np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]]), np.diag([1.0, -1.0]).astype(complex)]
    for _ in range(20):
        z = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        V, _ = np.linalg.qr(z)
        psi = rng.normal(size=(8, 2)) + 1j * rng.normal(size=(8, 2))
        lam = rng.normal(size=(8, 2)) + 1j * rng.normal(size=(8, 2))
        M = sum(np.outer(p, np.conj(l)) for p, l in zip(psi, lam))
        p0, p1, l0, l1 = psi[:, 0], psi[:, 1], lam[:, 0], lam[:, 1]
        h0 = np.sum(p0.imag * l0.real - p0.real * l0.imag)
        h1 = np.sum(p1.imag * l1.real - p1.real * l1.imag)
        h2 = np.sum(p0.imag * l1.real - p0.real * l1.imag + p1.imag * l0.real - p1.real * l0.imag)
        h3 = np.sum(p0.real * l1.real + p0.imag * l1.imag - p1.real * l0.real - p1.imag * l0.imag)
        for P in paulis:
            Q = V @ P @ V.conj().T
            full = np.trace(Q @ M).imag
            herm = Q[0, 0].real * h0 + Q[1, 1].real * h1 + Q[1, 0].real * h2 + Q[1, 0].imag * h3
            assert abs(full - herm) < 1e-9 * max(1.0, abs(full))
