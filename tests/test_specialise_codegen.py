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


@pytest.mark.parametrize("R,TB,lpos", [(5, 8, list(range(13))), (4, 8, [0, 1, 2, 3, 7, 8, 9, 10, 14, 15, 16, 18]),
                                       (5, 6, [0, 1, 2, 3, 5, 6, 9, 12, 13, 17, 19])])
def test_async_tile_loads_cover_the_tile_once(R, TB, lpos):
    """Every 16-byte slot of a buffer gets the amplitude pair the kernel's swizzled
    layout puts there: slot s = ii * 2^TB + tid holds tile elements swz_e(2s) and
    swz_e(2s) + 1, loaded from amplitude offset scatter(swz_e(2s)); over all loads of
    all threads each tile amplitude is loaded exactly once, into its own slot."""
    from qandle_amd._engine import jit

    K, T = R + TB, 1 << TB
    d_bit, d_inst = jit.async_load_offsets(lpos, R, TB)
    scat = lambda l: sum(1 << lpos[k] for k in range(K) if (l >> k) & 1)  # noqa: E731
    seen = set()
    for ii in range(1 << (R - 1)):
        for tid in range(T):
            off = d_inst[ii]
            for b in range(TB):
                if (tid >> b) & 1:
                    off ^= d_bit[b]
            s = ii * T + tid
            el = PL.swz(2 * s, 30)
            assert el % 2 == 0 and PL.swz(el, 30) == 2 * s       # the pair stays in order; swz is an involution
            assert off == scat(el) and off % 2 == 0               # 16-byte aligned pair start
            seen.add(off)
            seen.add(off + 1)
    assert seen == {scat(l) for l in range(1 << K)}


def test_async_pass_sources_compile(tmp_path):
    """The asynchronous-load pass kernels of a 14-qubit program compile (hipRTC, no
    device), and in the code object the loop's arithmetic starts while the next
    tile's loads are in flight: after each block of prefetch loads, a matrix-core
    instruction comes before any wait on the vector-memory counter."""
    import re
    import subprocess
    from qandle_amd import workloads
    from qandle_amd._engine import fused, jit

    w = workloads.rxryrz_ladder(n=14, B=2, depth=3, seed=4)
    circ = qandle.Circuit(w.layers, num_qubits=14)
    prog = fused.lower_circuit(circ.circuit)
    lead = fused.leading_rotations(prog)
    p = lead[0] if lead else prog
    pl = fused.get_plan(p, 5, 8, None, True)
    ws = fused.pass_words(p, pl, 2, "cpu")
    masks = fused.readout_fold(p)[1]
    off = 0
    srcs = []
    for i, (wd, _ns, lpos) in enumerate(ws.passes):
        words = wd.numpy()
        if i > 0:
            srcs.append(jit.pass_source_async(f"qs_pass_{i}", words, 14, 5, 8, lpos, i == len(ws.passes) - 1, 2,
                                              masks, "lds", off, 4, True))
        off += len(jit.lane_entries(words, 5, 8, False))
    assert srcs
    code = jit._compile(jit.PRELUDE + "\n\n".join(srcs), "async_probe")
    assert len(code) > 0
    objdump = "/opt/rocm/llvm/bin/llvm-objdump"
    if not os.path.exists(objdump):
        pytest.skip("llvm-objdump not installed")
    path = tmp_path / "async.hsaco"
    path.write_bytes(code)
    asm = subprocess.run([objdump, "-d", "--mcpu=gfx950", str(path)], capture_output=True, text=True).stdout
    ops = [m.group(1) for m in re.finditer(r"^\s+([a-z_0-9]+)", asm, re.M)]
    waits = [i for i, line in enumerate(asm.splitlines()) if "s_waitcnt" in line and "vmcnt" in line]
    lines = asm.splitlines()
    blocks = 0
    i = 0
    while i < len(lines):
        if "global_load_lds_dwordx4" in lines[i]:
            j = i
            while j < len(lines) and "global_load_lds_dwordx4" in lines[j]:
                j += 1
            # the first block of a kernel is its prologue load: nothing to overlap yet
            prev = "\n".join(lines[max(0, i - 40):i])
            if "s_barrier" in prev:
                nxt_wait = next((k for k in range(j, len(lines)) if "s_waitcnt" in lines[k] and "vmcnt" in lines[k]),
                                len(lines))
                nxt_mfma = next((k for k in range(j, len(lines)) if "v_mfma" in lines[k]), len(lines))
                assert nxt_mfma < nxt_wait, f"a vmcnt wait right after the prefetch at line {i}"
                blocks += 1
            i = j
        else:
            i += 1
    assert ops and waits and blocks >= 2


def test_direct_register_loads_and_stores_match_the_tile_layout():
    """Kernels that load the first sub-pass's registers straight from the state and
    store the last sub-pass's straight to it use, for register j of thread tq, the
    amplitude gbase | sum of 1 << GP[r] over j's bits (jit._register_offsets), where
    the tile path would use tile element gl ^ (xor of 1 << P[r]) and from there
    scatter(element): for the first and last sub-pass of every forward pass and
    adjoint step of a 20-qubit program the two addresses agree, and over all threads
    and registers each tile amplitude is moved exactly once."""
    from qandle_amd._engine import jit
    from qandle_amd import workloads
    from qandle_amd._engine import fused

    n = 20
    w = workloads.rxryrz_ladder(n=n, B=2, depth=4, seed=3)
    circ = qandle.Circuit(w.layers, num_qubits=n)
    prog = fused.lower_circuit(circ.circuit)
    lead = fused.leading_rotations(prog)
    p = lead[0] if lead else prog
    cases = []
    R, TB = 5, 8
    pl = fused.get_plan(p, R, TB, None, True)
    for wd, _ns, lpos in fused.pass_words(p, pl, 2, "cpu").passes:
        cases.append((R, TB, lpos, PL.decode_pass(wd.numpy(), R, TB)))
    # the adjoint steps (gradient tile shape) use the same register layout
    R, TB = fused.choose_shape(n, True)
    pl = fused.get_plan(prog, R, TB, fused.GRAD_COST, True)
    ws, _nmain = fused.adjoint_words(prog, pl, 2, "cpu", True, fused.GRAD_COST)
    for wd, _ns, lpos in ws.passes:
        cases.append((R, TB, lpos, PL.decode_adjoint_pass(wd.numpy(), R, TB)))
    for R, TB, lpos, decoded in cases:
        K = R + TB
        lpos = [int(x) for x in lpos]
        scat = lambda l: sum(1 << lpos[k] for k in range(K) if (l >> k) & 1)  # noqa: E731
        for sub in (decoded[0], decoded[-1]):
            roff, r0 = jit._register_offsets(sub, R)
            assert r0 is None or sub.GP[r0] == 0
            seen = []
            for tq in range(1 << TB):
                gl = sum(((tq >> i) & 1) << sub.G[i] for i in range(TB))
                gbase = sum(((tq >> i) & 1) << sub.GG[i] for i in range(TB))
                for j in range(1 << R):
                    el = gl
                    for r in range(R):
                        if (j >> r) & 1:
                            el ^= 1 << sub.P[r]
                    assert gbase & roff[j] == 0 and gbase | roff[j] == scat(el)
                    seen.append(gbase | roff[j])
            assert sorted(seen) == sorted(scat(l) for l in range(1 << K))
            # lane bits: the lowest tile bits that are neither registers nor wave bits, ascending
            taken = set(sub.P[:R]) | set(sub.G[6:TB])
            assert [sub.G[i] for i in range(6)] == [b for b in range(K) if b not in taken][:6]
