"""Serving loop: admission by HBM footprint (CPU, host logic only) and, on the
GPU, identical results alone and in a concurrent mix."""
import numpy as np
import pytest
import torch

import qandle_amd as qandle
from oracle import fp64
from qandle_amd import workloads
from qandle_amd._engine import serving


def _req(rid, n, B, depth, seed, device="cpu", arrival=0.0):
    gen = torch.Generator().manual_seed(seed)
    w = workloads.hardware_efficient(n, B, depth, gen)
    circ = qandle.Circuit(w.layers, num_qubits=n)
    if device != "cpu":
        circ = circ.to(device)
    return serving.Request(rid, circ, {"x": w.inputs["x"].to(device)}, arrival=arrival), w


def test_admission_footprint_chunking_and_rejection():
    eng = _host_engine(10 * (2 ** 12) * 8 + 50_000)  # host-side parts only
    r, _ = _req(0, 12, 37, 2, 1)
    eng.prepare(r)
    assert r.error is None and r.n == 12 and r.B == 37
    assert r.footprint > eng.cap
    assert r.chunks[0] == (0, r.chunks[0][1]) and r.chunks[-1][1] == 37
    sizes = [hi - lo for lo, hi in r.chunks]
    assert all(serving.state_footprint(12, s, 5, 6) <= eng.cap for s in sizes)
    big, _ = _req(1, 16, 1, 1, 2)
    eng.cap = 2 ** 16 * 8 // 2
    eng.prepare(big)
    assert big.error is not None and "admission cap" in big.error


def test_poisson_stream_is_seeded():
    a = serving.poisson_requests.__wrapped__ if hasattr(serving.poisson_requests, "__wrapped__") else None
    assert a is None  # plain function
    rng1 = np.random.default_rng(3).exponential(1.0, 5)
    rng2 = np.random.default_rng(3).exponential(1.0, 5)
    assert np.array_equal(rng1, rng2)


@pytest.mark.gpu
def test_results_identical_alone_and_in_mix(device):
    reqs_mix, reqs_alone, ws = [], [], []
    spec = [(8, 5, 3), (12, 3, 4), (14, 2, 2), (10, 7, 5), (13, 4, 3), (9, 1, 2), (15, 2, 4), (11, 6, 3)] * 3
    for i, (n, B, d) in enumerate(spec):
        r, w = _req(i, n, B, d, 100 + i, device, arrival=0.0005 * i)
        reqs_mix.append(r)
        ws.append(w)
        r2, _ = _req(i, n, B, d, 100 + i, device, arrival=0.0)
        reqs_alone.append(r2)
    # requests of 12 qubits and more run compiled kernels, built ahead (in this process: a
    # process that already uses the GPU starts no compile workers)
    nstruct = serving.precompile(reqs_mix, 12, workers=1)
    assert nstruct == len({(n, d) for n, _, d in spec if n >= 12})
    eng = serving.ServingEngine(device, cap_bytes=64 * 2 ** 20, streams=4, compiled_from_qubits=12)
    for r in reqs_mix:
        eng.prepare(r)
    assert {r.compiled for r in reqs_mix} == {True, False}
    done = eng.serve(reqs_mix)
    assert len(done) == len(spec) and all(r.error is None for r in done)
    assert eng.peak_resident <= 64 * 2 ** 20
    for r in reqs_alone:
        solo = serving.ServingEngine(device, cap_bytes=64 * 2 ** 20, streams=1, max_inflight=1,
                                     compiled_from_qubits=12)
        solo.prepare(r)
        solo.serve([r])
    for a, b, w in zip(reqs_mix, reqs_alone, ws):
        assert torch.equal(a.result, b.result), f"request {a.rid} differs alone vs in the mix"
    for i in (0, 3, 6):
        w = ws[i]
        ref = fp64.measure_prob(fp64.simulate(w.n, w.oracle_ops()), w.n)
        assert np.abs(reqs_mix[i].result.cpu().numpy() - ref).max() < 1e-5
    st = serving.latency_stats(done)
    assert st["p50_ms"] > 0 and st["p99_ms"] >= st["p50_ms"]
    assert 0 < st["service_p50_ms"] <= st["p99_ms"]


class _DoneEvent:
    def query(self):
        return True


def _host_engine(cap, small_work=-1):
    """A ServingEngine with its host-side state only (no streams, no device); with the
    default ``small_work`` every chunk counts as large."""
    import time

    eng = serving.ServingEngine.__new__(serving.ServingEngine)
    eng.cap, eng.max_inflight, eng.max_wait_s, eng.compiled_from = int(cap), 64, 2.0, None
    eng.warm = False  # no device: nothing is encoded or uploaded ahead
    eng.max_chunk_frac, eng.small_work, eng.big_cap = 0.25, small_work, int(cap * 0.9)
    eng.stream_work = [0] * 4
    eng.queue, eng.inflight, eng.done = [], [], []
    eng.resident = eng.small_resident = eng.peak_resident = 0
    eng.t0 = time.perf_counter()
    return eng


def test_oom_rejection_drops_every_chunk_of_the_request(monkeypatch):
    """An allocator failure with nothing in flight rejects the whole request once:
    its other queued chunks never launch and the loop keeps serving the rest."""
    eng = _host_engine(3 * (2 ** 12) * 8 + 60_000)
    big, _ = _req(0, 12, 10, 2, 1)
    small, _ = _req(1, 10, 2, 2, 2)
    eng.prepare(big)
    eng.prepare(small)
    assert len(big.chunks) >= 3 and len(small.chunks) == 1
    launched = []

    def fake_launch(ch):
        if ch.req is big:
            raise torch.cuda.OutOfMemoryError("fake allocator failure")
        launched.append(ch)
        ch.out, ch.event = torch.full((ch.hi - ch.lo, ch.req.n), 0.5), _DoneEvent()
        eng._hold(ch)

    monkeypatch.setattr(eng, "_launch", fake_launch)
    eng.submit(big)
    eng.submit(small)
    while eng.inflight:
        eng.poll()
    assert [r.rid for r in eng.done].count(0) == 1 and [r.rid for r in eng.done].count(1) == 1
    assert big.error is not None and big.result is None and big.chunks == []
    assert not any(c.req is big for c in eng.queue) and not any(c.req is big for c in launched)
    assert small.error is None and small.result.shape == (2, 10)
    assert eng.resident == 0 and not eng.queue


def test_chunk_of_a_rejected_request_is_dropped_when_it_retires():
    eng = _host_engine(1e12)
    r, _ = _req(0, 10, 4, 2, 3)
    eng.prepare(r)
    ch = serving._Chunk(r, 0, 4, 1000, event=_DoneEvent(), out=torch.zeros(4, 10))
    eng.inflight.append(ch)
    eng.resident = 1000
    eng._reject(r, "rejected while running")
    eng._reject(r, "a second failure is not reported twice")
    eng.poll()
    assert eng.done == [r] and r.result is None and eng.resident == 0 and not eng.inflight


def test_admission_order_shortest_first_aging_and_backfill(monkeypatch):
    """Young chunks go smallest first; a chunk that waited past max_wait_s goes
    first, and when it does not fit, only chunks that leave room for it are
    admitted until it does."""
    eng = _host_engine(1000)
    launched = []

    def fake_launch(ch):
        launched.append(ch.req.rid)
        ch.event, ch.out = _Pending(), torch.zeros(ch.hi - ch.lo, 1)
        eng._hold(ch)

    monkeypatch.setattr(eng, "_launch", fake_launch)

    def req(rid, nbytes, queued):
        r = serving.Request(rid, None, {})
        r.n, r.B, r.t_queued, r.chunks = 1, 1, queued, [(0, 1)]
        return serving._Chunk(r, 0, 1, nbytes)

    now = eng.now()
    # 600 B already resident; queue: a young 300, a young 100, an aged 500 (waited 5 s)
    eng._hold(serving._Chunk(serving.Request(99, None, {}), 0, 1, 600, event=_Pending()))
    eng.queue = [req(1, 300, now), req(2, 100, now), req(3, 500, now - 5.0)]
    eng._admit()
    # the aged 500 does not fit (600 + 500 > 1000) and reserves 500: only chunks that keep
    # resident + bytes <= 500 may pass, and 600 is already resident, so nothing is admitted
    assert launched == []
    # the resident work retires: the aged chunk goes first, then the smallest young ones that fit
    eng.inflight, eng.resident, eng.small_resident = [], 0, 0
    eng._admit()
    assert launched == [3, 2, 1][: len(launched)] and launched[0] == 3 and launched[1] == 2
    assert eng.resident <= 1000


class _Pending:
    def query(self):
        return False


def test_small_pool_and_stream_choice(monkeypatch):
    """Small chunks (few amplitude updates) may use the pool that large chunks
    never take even while an aged large chunk holds a reservation; large chunks
    stay within their share; each large chunk goes to the stream with the least
    outstanding work, small ones to stream 0.  Chunks of a large request are
    bounded to a quarter of the cap."""
    eng = _host_engine(1000, small_work=50)
    launched = []

    def fake_launch(ch):
        ch.stream = eng.pick_stream(ch)
        launched.append((ch.req.rid, ch.stream))
        ch.event, ch.out = _Pending(), torch.zeros(ch.hi - ch.lo, 1)
        eng._hold(ch)

    monkeypatch.setattr(eng, "_launch", fake_launch)

    def chunk(rid, nbytes, work, queued):
        r = serving.Request(rid, None, {})
        r.n, r.B, r.t_queued, r.chunks = 1, 1, queued, [(0, 1)]
        return serving._Chunk(r, 0, 1, nbytes, work=work)

    now = eng.now()
    # large work of 850 B in flight (big share 900): an aged large 200 reserves; a young large
    # 40 would cross the share, a young small 60 fits the small pool (1000 - 900) regardless
    for b, w in ((400, 1000), (450, 500)):
        fake_launch(chunk(90 + b, b, w, now))
    assert [s for _r, s in launched] == [1, 2]
    launched.clear()
    eng.queue = [chunk(1, 200, 10 ** 6, now - 5.0), chunk(2, 40, 100, now), chunk(3, 60, 10, now)]
    eng._admit()
    assert launched == [(3, 0)]
    assert eng.small_resident == 60 and eng.resident == 910
    # the first large chunk retires: the aged one goes to the emptied stream 1 (streams 1 and 3
    # hold no work; ties go to the lower index), the young large one then to stream 3
    first = eng.inflight[0]
    eng.inflight.remove(first)
    eng._release(first)
    eng._admit()
    assert launched[1:] == [(1, 1), (2, 3)]
    assert eng.resident - eng.small_resident <= eng.big_cap
    # chunk bound: a request of 37 samples of 12 qubits under a cap of 10 samples' bytes
    eng.cap = 10 * (2 ** 12) * 8 + 50_000
    r, _ = _req(0, 12, 37, 2, 1)
    eng.prepare(r)
    sizes = [hi - lo for lo, hi in r.chunks]
    assert all(serving.state_footprint(12, s, 5, 6) <= max(eng.cap // 4, serving.state_footprint(12, 1, 5, 6))
               for s in sizes)
    assert r.chunks[-1][1] == 37 and len(r.chunks) >= 4


def test_precompile_workers_fill_the_code_cache(tmp_path, monkeypatch):
    """precompile in worker processes (hipRTC only, nothing loaded onto a device)
    writes one code object per program structure; building the same structures
    again in this process then finds every source in the cache."""
    import os

    from qandle_amd._engine import jit

    cache = str(tmp_path / "jit")
    monkeypatch.setenv("QANDLE_AMD_JIT_CACHE", cache)  # read by the spawned workers at import
    monkeypatch.setattr(jit, "CACHE_DIR", cache)
    reqs = [_req(i, n, 2, d, 40 + i)[0] for i, (n, d) in enumerate([(13, 2), (14, 3), (13, 2), (9, 2)])]
    assert serving.precompile(reqs, 12, workers=2) == 2   # (13, 2) twice; 9 qubits is below the threshold
    assert len([f for f in os.listdir(cache) if f.endswith(".hsaco")]) == 2

    def no_compile(*a, **k):
        raise AssertionError("source not found in the code-object cache")

    monkeypatch.setattr(jit, "_compile", no_compile)
    assert serving.precompile(reqs, 12, workers=1) == 2


def test_state_arena_regions():
    """The serving arena's host bookkeeping: aligned regions, small ones from the
    top end and large ones from the bottom, coalescing on return, and no region
    handed out twice (checked against a byte map)."""
    import numpy as np

    A = serving.StateArena.__new__(serving.StateArena)
    A.size, A.free = 64 * serving.StateArena.ALIGN, [(0, 64 * serving.StateArena.ALIGN)]
    al = serving.StateArena.ALIGN
    used = np.zeros(64, dtype=bool)
    rng = np.random.default_rng(0)
    live = []
    for step in range(400):
        if live and (rng.random() < 0.45 or not A.fits(al)):
            off, nb = live.pop(int(rng.integers(len(live))))
            A.give(off, nb)
            used[off // al: off // al + -(-nb // al)] = False
        else:
            nb = int(rng.integers(1, 9 * al))
            top = bool(rng.random() < 0.5)
            off = A.take(nb, top=top)
            if off is None:
                assert not A.fits(nb)
                continue
            blocks = -(-nb // al)
            assert off % al == 0 and off + blocks * al <= A.size
            assert not used[off // al: off // al + blocks].any()
            used[off // al: off // al + blocks] = True
            live.append((off, nb))
        # free list: sorted, disjoint, coalesced, and exactly the unused blocks
        offs = [o for o, _ in A.free]
        assert offs == sorted(offs)
        for (o1, l1), (o2, _l2) in zip(A.free, A.free[1:]):
            assert o1 + l1 < o2
        free_map = np.zeros(64, dtype=bool)
        for o, ln in A.free:
            free_map[o // al: (o + ln) // al] = True
        assert (free_map == ~used).all()
    for off, nb in live:
        A.give(off, nb)
    assert A.free == [(0, A.size)]
    # small from the top, large from the bottom
    a = A.take(al, top=True)
    b = A.take(2 * al)
    assert a == A.size - al and b == 0


@pytest.mark.gpu
def test_prepared_launches_upload_nothing(device, monkeypatch):
    """After prepare, a chunk's launch does no host-to-device upload (an upload
    would wait behind the kernels queued on a busy stream): every plan, word
    set, table and kernel it needs was built for its chunk size ahead.  The
    requests are chunked (a cap that holds two samples of the largest), mix
    compiled and interpreting kernels, and match the same requests served
    without the arena."""
    from qandle_amd._engine import lib as L

    spec = [(11, 5, 2, 1), (13, 6, 3, 2), (15, 4, 2, 3)]
    per = serving.state_footprint(15, 1, *serving.kernel_shape(15, True))
    reqs = [_req(i, n, B, d, s, device)[0] for i, (n, B, d, s) in enumerate(spec)]
    plain = [_req(i, n, B, d, s, device)[0] for i, (n, B, d, s) in enumerate(spec)]
    eng = serving.ServingEngine(device, cap_bytes=9 * per, streams=2, compiled_from_qubits=13)
    for r in reqs:
        eng.prepare(r)
    assert any(len(r.chunks) > 1 for r in reqs) and {r.compiled for r in reqs} == {True, False}

    def no_upload(*a, **k):
        raise AssertionError("a prepared launch uploaded host data")

    monkeypatch.setattr(L, "upload", no_upload)
    done = eng.serve(reqs)
    monkeypatch.undo()
    assert all(r.error is None for r in done)
    ref = serving.ServingEngine(device, cap_bytes=9 * per, streams=2, compiled_from_qubits=13, arena=False)
    for r in plain:
        ref.prepare(r)
    ref.serve(plain)
    for a, b in zip(reqs, plain):
        assert torch.equal(a.result, b.result)
