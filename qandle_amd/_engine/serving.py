"""Continuous-batching serving loop for circuit evaluations on one GPU.

The reference has no serving component; this is the request path a model
server puts in front of ``Circuit.forward``:

* Requests (a lowered circuit, its inputs, an arrival time) wait in a queue.
* Admission is by HBM footprint: a chunk of a request is admitted when the
  resident bytes of everything in flight plus its own state (B x 2^n x 8 B,
  plus its measurement partials and factors) stay under the cap and an
  in-flight slot is free.  A request whose batch alone exceeds the cap is
  split into batch chunks that each fit; a single sample that does not fit is
  rejected with an error.
* Chunks are also bounded to ``max_chunk_frac`` of the cap, so that a large
  request runs as several chunks that leave room beside them.
* Order: the smallest waiting chunk first (shortest-footprint first keeps small
  requests from queueing behind large ones), except that chunks which have
  waited longer than ``max_wait_s`` go first, in arrival order.  When such an
  aged chunk does not fit yet, memory is reserved for it: other chunks are only
  admitted while they leave room for it (backfilling), so it is admitted as
  soon as enough in-flight work retires and nothing starves.
* Small and large work: a chunk of at most ``small_work`` amplitude updates
  (milliseconds of device time) is small.  Large chunks may hold at most
  ``big_share`` of the cap; the rest is a pool that small chunks can always
  use, an aged reservation notwithstanding, so a burst of large requests does
  not hold up the many small ones.
* Admitted work is enqueued at once on a HIP stream (default priority) and
  overlaps on the device with everything else in flight.  A stream runs its
  work in order, so small chunks get a stream of their own and each large chunk
  goes to the other stream with the least outstanding work: a small request
  never waits behind a large one in a stream.  The host never blocks on a
  request: completion is detected by polling a HIP event, after which the
  request's memory is released and more requests are admitted (continuous
  batching, not a queue drained one batch at a time).
* Kernels: requests of at least ``compiled_from_qubits`` qubits run the pass
  kernels compiled for their program structure (``precompile`` builds them
  ahead of serving, in worker processes, into the code-object cache); smaller
  requests, which are bound by launch overhead rather than by the kernels,
  run the library's interpreting kernels and never wait for a compiler.  The
  choice is a function of the request alone.
* Results are a pure function of (circuit, inputs): the kernel choice, shape
  and fusion plan do not depend on load or chunking, and each sample's tiles
  are reduced in a fixed order, so a request computes the same bits alone, in
  a mix, and split into chunks.
* Latency = completion time - arrival time (host clock); service time =
  completion - first admission.  p50/p99 of both are reported with throughput.
"""
from __future__ import annotations

import dataclasses
import multiprocessing as mp
import threading
import time
import typing

import numpy as np
import torch

from . import fused
from . import lib as L


@dataclasses.dataclass
class Request:
    rid: int
    circuit: typing.Any            # built qandle_amd.Circuit (measured at the end)
    inputs: typing.Dict[str, torch.Tensor]
    arrival: float = 0.0           # seconds after the serving clock starts
    # filled by the engine
    n: int = 0
    B: int = 0
    footprint: int = 0
    work: int = 0                  # amplitude updates per sample
    compiled: bool = False
    chunks: list = dataclasses.field(default_factory=list)
    parts: dict = dataclasses.field(default_factory=dict)
    result: typing.Optional[torch.Tensor] = None
    error: typing.Optional[str] = None
    t_queued: float = -1.0
    t_admit: float = -1.0
    t_done: float = -1.0

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival

    @property
    def service(self) -> float:
        return self.t_done - self.t_admit


@dataclasses.dataclass
class _Chunk:
    req: Request
    lo: int
    hi: int
    bytes: int
    event: typing.Any = None
    out: typing.Any = None
    keep: typing.Any = None
    work: int = 0       # amplitude updates: (gates + qubits) x samples x 2^n
    stream: int = -1    # index of the stream it was enqueued on
    region: typing.Any = None  # (offset, bytes) of its state in the engine's arena

    @property
    def state_bytes(self) -> int:
        """Bytes of its (samples, 2^n) complex64 state."""
        return ((self.hi - self.lo) << self.req.n) * 8


def state_footprint(n: int, B: int, R: int, TB: int) -> int:
    """HBM bytes a running evaluation of (n, B) holds: the state plus partials
    (either partial layout) and factor tables."""
    lib = L.load()
    part = max(lib.qs_fused_measure_work_bytes(B, n, R, TB), lib.qs_readout_work_bytes(B, n, R, TB), 0)
    return B * (2 ** n) * 8 + part + B * n * 32 + 4096


class StateArena:
    """One device allocation that the serving loop carves its chunks' states
    from, so that taking and returning a state is host bookkeeping.  With
    states of KBs to tens of GB coming and going, a caching allocator that
    cannot place a large one frees its cached blocks first, which waits for
    every stream; the loop then stopped (no admissions, no completions) for as
    long as the largest kernels in flight ran.  Regions are aligned to 64 KiB;
    small states are placed from the top end and large ones from the bottom,
    so that small ones do not split the space large ones need."""

    ALIGN = 1 << 16

    def __init__(self, nbytes: int, device):
        self.size = max(int(nbytes) // self.ALIGN * self.ALIGN, self.ALIGN)
        self.buf = torch.empty(self.size // 8, dtype=torch.complex64, device=device)
        self.free = [(0, self.size)]  # (offset, length), sorted by offset, neighbours coalesced

    def _span(self, nbytes: int) -> int:
        return -(-int(nbytes) // self.ALIGN) * self.ALIGN

    def fits(self, nbytes: int) -> bool:
        need = self._span(nbytes)
        return any(ln >= need for _, ln in self.free)

    def take(self, nbytes: int, top: bool = False):
        """Offset of a region of ``nbytes`` (None when no free extent holds it):
        the lowest extent's start, or with ``top`` the highest extent's end."""
        need = self._span(nbytes)
        order = range(len(self.free) - 1, -1, -1) if top else range(len(self.free))
        for i in order:
            off, ln = self.free[i]
            if ln < need:
                continue
            if top:
                self.free[i] = (off, ln - need)
                got = off + ln - need
            else:
                self.free[i] = (off + need, ln - need)
                got = off
            if self.free[i][1] == 0:
                del self.free[i]
            return got
        return None

    def give(self, off: int, nbytes: int):
        """Returns a region taken with the same size."""
        need = self._span(nbytes)
        lo = 0
        while lo < len(self.free) and self.free[lo][0] < off:
            lo += 1
        self.free.insert(lo, (off, need))
        if lo + 1 < len(self.free) and off + need == self.free[lo + 1][0]:
            self.free[lo] = (off, need + self.free[lo + 1][1])
            del self.free[lo + 1]
        if lo > 0 and self.free[lo - 1][0] + self.free[lo - 1][1] == off:
            self.free[lo - 1] = (self.free[lo - 1][0], self.free[lo - 1][1] + self.free[lo][1])
            del self.free[lo]

    def state(self, off: int, B: int, n: int) -> torch.Tensor:
        """The (B, 2^n) complex64 state at byte offset ``off``."""
        return self.buf[off // 8: off // 8 + B * (1 << n)].view(B, 1 << n)


def kernel_shape(n: int, compiled: bool):
    """(R, TB) of a request's passes: the compiled forward shape, or the
    interpreting kernels' one-wave tiles."""
    return fused.choose_shape(n) if compiled else fused.interpreted_shape(n)


def _structure(prog):
    """Hashable structure of a lowered program: what its compiled kernels depend on."""
    gates = tuple((g.kind, g.pos, g.ctrl) for g in prog.gates)
    return prog.n, gates, tuple(map(str, prog.init[:3])), prog.measure


def _spec(prog):
    """(n, gate spec, embed spec) of a lowered program, for fused.program_from_spec
    (angles are left out: the kernels do not depend on them)."""
    names = {fused.PL.RX: "rx", fused.PL.RY: "ry", fused.PL.RZ: "rz"}
    n = prog.n
    gates = [("cnot", (n - 1 - g.ctrl, n - 1 - g.pos)) if g.kind == fused.PL.CNOT
             else (names[g.kind], n - 1 - g.pos, 0.0) for g in prog.gates]
    embed = None
    if prog.init[0] == "embed":
        embed = ("xyz"[prog.init[1]], [n - 1 - b for b in prog.init[2]], prog.init[3])
    return n, gates, embed


def _compile_worker(specs):
    """Compile (into the on-disk code-object cache; nothing is loaded onto a
    device) the forward kernels of each (n, gate spec, embed spec) program."""
    for n, gates, embed in specs:
        prog = fused.program_from_spec(n, gates, embed)
        R, TB = kernel_shape(n, True)
        fused.forward_kernels(prog, 1, R, TB, fused.DEFAULT_COST, "cpu", load=False)
    return len(specs)


def precompile(requests, compiled_from_qubits: int, workers: int = 8) -> int:
    """Compile the pass kernels of every distinct program structure among the
    requests that will run compiled kernels, in ``workers`` processes (hipRTC
    compiles one program at a time per process).  Angles are data, so one
    structure's kernels serve every request of that structure.  Returns the
    number of structures."""
    seen, specs = set(), []
    for r in requests:
        prog = fused.lower_circuit(r.circuit.circuit)
        if prog is None or prog.n < compiled_from_qubits:
            continue
        key = _structure(prog)
        if key not in seen:
            seen.add(key)
            specs.append(_spec(prog))
    if not specs:
        return 0
    workers = max(1, min(workers, len(specs)))
    # largest programs first, dealt round-robin so the workers finish together
    specs.sort(key=lambda s: -(len(s[1]) << s[0]))
    groups = [specs[i::workers] for i in range(workers)]
    if workers == 1:
        _compile_worker(groups[0])
    else:
        with mp.get_context("spawn").Pool(workers) as pool:
            pool.map(_compile_worker, groups)
    return len(specs)


_STREAMS: dict = {}


def serving_streams(device, count: int):
    """The serving streams of a device: created once per process and shared by
    every engine on it.  HIP maps streams onto a fixed number of hardware
    queues (GPU_MAX_HW_QUEUES, 4 on the machines measured), and a kernel
    waits in order behind every kernel queued before it on its hardware
    queue.  Engines that each took new streams from torch's stream pool put
    the small-request stream on a queue shared with a large-chunk stream often
    enough to move the median latency at 70% load from 3.5-4.5 ms to 120 ms
    (measured on the serving benchmark's request mix)."""
    dev = torch.device(device)
    key = (dev.index if dev.index is not None else torch.cuda.current_device(), int(count))
    with _STREAMS_LOCK:
        got = _STREAMS.get(key)
        if got is None:
            got = _STREAMS[key] = [torch.cuda.Stream(dev) for _ in range(count)]
    return got


_STREAMS_LOCK = threading.Lock()


class ServingEngine:
    def __init__(self, device, cap_bytes: float = 200e9, streams: int = 4, max_inflight: int = 64,
                 max_wait_s: float = 2.0, compiled_from_qubits: typing.Optional[int] = None,
                 max_chunk_frac: float = 0.25, small_work: float = 1e11, big_share: float = 0.9,
                 warm: bool = True, arena: bool = True):
        self.device = torch.device(device)
        self.warm = warm
        self.use_arena = arena
        self.arena = None
        self.max_wait_s = max_wait_s
        self.cap = int(cap_bytes)
        self.compiled_from = compiled_from_qubits
        self.max_chunk_frac = max_chunk_frac
        self.small_work = small_work
        self.big_cap = int(self.cap * big_share)
        self.streams = serving_streams(self.device, streams)
        self.stream_work = [0] * streams
        self.max_inflight = max_inflight
        self.queue: typing.List[_Chunk] = []
        self.inflight: typing.List[_Chunk] = []
        self.resident = 0
        self.small_resident = 0
        self.peak_resident = 0
        self.done: typing.List[Request] = []
        self.t0 = None
        L.load()

    # ----------------------------------------------------------- admission
    def prepare(self, req: Request):
        """Lower, plan, encode and upload (and, for compiled requests, load their
        kernels) before serving; split over-cap batches into chunks.  Everything
        a chunk's launch needs from the host is built here, for every chunk
        size, so that a launch while the streams are busy never waits for them
        (an upload or a copy back would wait behind the queued kernels)."""
        prog = fused.lower_circuit(req.circuit.circuit)
        if prog is None or not prog.measure:
            raise ValueError("serving takes circuits that end in MeasureProbability")
        x = req.inputs[prog.init[3]] if prog.init[0] == "embed" else None
        req.n = prog.n
        req.B = 1 if x is None or x.dim() == 1 else x.shape[0]
        req.compiled = self.compiled_from is not None and prog.n >= self.compiled_from
        R, TB = kernel_shape(prog.n, req.compiled)
        per = state_footprint(req.n, 1, R, TB)
        req.footprint = state_footprint(req.n, req.B, R, TB)
        req.work = (len(prog.gates) + prog.n) << prog.n  # amplitude updates per sample
        if per > self.cap:
            req.error = f"one sample needs {per} bytes, over the {self.cap}-byte admission cap"
            return req
        limit = max(per, int(self.cap * self.max_chunk_frac))
        step = req.B if req.footprint <= limit else max(1, limit // per)
        while step > 1 and state_footprint(req.n, step, R, TB) > limit:
            step -= 1
        req.chunks = [(lo, min(req.B, lo + step)) for lo in range(0, req.B, step)]
        if self.warm:
            cm = fused.DEFAULT_COST if req.compiled else fused.INTERPRETED_COST
            for b in sorted({hi - lo for lo, hi in req.chunks}):
                fused.warm(prog, b, R, TB, cm, req.compiled, self.device)
            fused.param_angles(prog, self.device)
        return req

    def submit(self, req: Request):
        if req.error:
            req.t_admit = req.t_done = self.now()
            self.done.append(req)
            return
        req.t_queued = self.now()
        R, TB = kernel_shape(req.n, req.compiled)
        for lo, hi in req.chunks:
            self.queue.append(_Chunk(req, lo, hi, state_footprint(req.n, hi - lo, R, TB), work=req.work * (hi - lo)))
        self._admit()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def _arena(self):
        """The state arena (the admission cap's bytes, allocated on first use;
        serve() drops it at its end, and the caching allocator keeps the block
        for the next arena of the same size)."""
        if not getattr(self, "use_arena", False):
            return None
        if self.arena is None:
            self.arena = StateArena(self.cap, self.device)
        return self.arena

    def _order(self):
        """Aged chunks (waited longer than max_wait_s) in arrival order, then the
        rest by footprint (ties by arrival)."""
        now = self.now()
        aged = [c for c in self.queue if now - c.req.t_queued > self.max_wait_s]
        young = [c for c in self.queue if now - c.req.t_queued <= self.max_wait_s]
        young.sort(key=lambda c: (c.bytes, c.req.t_queued, c.lo))
        return aged + young, len(aged)

    def small(self, ch: _Chunk) -> bool:
        return ch.work <= self.small_work

    def fits(self, ch: _Chunk, reserve: int) -> bool:
        """Whether ``ch`` may be admitted now, with ``reserve`` bytes held back for an
        aged chunk: large chunks stay within their share of the cap; small ones fit
        beside the reservation or in the small pool (the cap less the large share,
        which large chunks never take)."""
        if self.resident + ch.bytes > self.cap:
            return False
        arena = self._arena()
        if arena is not None and not arena.fits(ch.state_bytes):
            return False
        if self.small(ch):
            return (self.resident + ch.bytes + reserve <= self.cap
                    or self.small_resident + ch.bytes <= self.cap - self.big_cap)
        big = self.resident - self.small_resident
        return big + ch.bytes + reserve <= self.big_cap and self.resident + ch.bytes + reserve <= self.cap

    def _admit(self):
        """Admit in _order while memory and slots allow; the first aged chunk that
        does not fit reserves its bytes, and later chunks only backfill the rest."""
        while self.queue and len(self.inflight) < self.max_inflight:
            order, naged = self._order()
            reserve, pick = 0, None
            for k, ch in enumerate(order):
                if self.fits(ch, reserve) or not self.inflight:
                    pick = ch
                    break
                if reserve == 0 and k < naged:
                    reserve = ch.bytes
            if pick is None:
                return
            self.queue.remove(pick)
            try:
                self._launch(pick)
            except torch.cuda.OutOfMemoryError:
                # the allocator could not place it after all: retry once memory frees up,
                # or reject it when nothing else is holding memory
                if self.inflight:
                    self.queue.append(pick)
                    return
                self._reject(pick.req, "out of device memory at admission")

    def _reject(self, req: Request, why: str):
        """Fail a request once: its queued chunks leave the queue, chunks of it still
        in flight are dropped when they retire (poll), and it is reported done now."""
        if req.error is not None:
            return
        req.error = why
        self.queue = [c for c in self.queue if c.req is not req]
        req.chunks = []
        req.parts = {}
        req.t_done = self.now()
        self.done.append(req)

    def pick_stream(self, ch: _Chunk) -> int:
        """Stream 0 takes the small chunks; a large chunk goes to the stream (1 and
        up) with the least outstanding work (with one stream, everything is on it)."""
        if len(self.stream_work) == 1 or self.small(ch):
            return 0
        return min(range(1, len(self.stream_work)), key=lambda i: (self.stream_work[i], i))

    def _launch(self, ch: _Chunk):
        req = ch.req
        ch.stream = self.pick_stream(ch)
        s = self.streams[ch.stream]
        s.wait_stream(torch.cuda.current_stream(self.device))
        arena = self._arena()
        if arena is not None:
            off = arena.take(ch.state_bytes, top=self.small(ch))
            if off is None:  # fits() checked an extent; nothing between took it
                raise RuntimeError("state arena has no extent for an admitted chunk")
            ch.region = (off, ch.state_bytes)
            psi = arena.state(off, ch.hi - ch.lo, req.n)
        else:
            # from the allocating stream's pool, recorded on the launch stream: one pool serves
            # the chunks of every stream instead of each stream caching its own blocks
            psi = torch.empty(ch.hi - ch.lo, 2 ** req.n, dtype=torch.complex64, device=self.device)
            psi.record_stream(s)
        try:
            self._run(ch, s, psi)
        except BaseException:
            if ch.region is not None:
                s.synchronize()  # kernels enqueued before the failure may still write the region
            self._give_region(ch)
            raise
        if req.t_admit < 0:
            req.t_admit = self.now()
        self._hold(ch)

    def _give_region(self, ch: _Chunk):
        if ch.region is not None and self.arena is not None:
            self.arena.give(*ch.region)
        ch.region = None

    def _run(self, ch: _Chunk, s, psi):
        req = ch.req
        with torch.cuda.stream(s), torch.no_grad():
            kw = dict(req.inputs)
            prog = fused.lower_circuit(req.circuit.circuit)
            if prog.init[0] == "embed":
                x = kw[prog.init[3]]
                kw[prog.init[3]] = x[ch.lo:ch.hi] if x.dim() == 2 else x
            inp = fused.gather_inputs(prog, req.circuit.circuit, None, kw)
            R, TB = kernel_shape(prog.n, req.compiled)
            cs = fused.build_coeffs(prog, inp)
            efac = fused.product_factors(prog, inp)
            cm = fused.DEFAULT_COST if req.compiled else fused.INTERPRETED_COST
            _, probs = fused.execute(prog, inp, cs, efac, R, TB, cm=cm, want_state=False, specialise=req.compiled,
                                     out=psi)
            ch.out = probs
            ch.keep = (cs, efac)
            ev = torch.cuda.Event()
            ev.record(s)
            ch.event = ev

    def _hold(self, ch: _Chunk):
        """Account an admitted chunk's memory and stream work."""
        self.resident += ch.bytes
        if self.small(ch):
            self.small_resident += ch.bytes
        if ch.stream >= 0:
            self.stream_work[ch.stream] += ch.work
        self.peak_resident = max(self.peak_resident, self.resident)
        self.inflight.append(ch)

    def _release(self, ch: _Chunk):
        self._give_region(ch)  # its kernels have run (its event completed)
        self.resident -= ch.bytes
        if self.small(ch):
            self.small_resident -= ch.bytes
        if ch.stream >= 0:
            self.stream_work[ch.stream] -= ch.work

    def poll(self) -> int:
        """Retire finished chunks, release their memory, admit more.  Returns #retired."""
        still, retired = [], 0
        for ch in self.inflight:
            if ch.event.query():
                retired += 1
                self._release(ch)
                req = ch.req
                out, ch.out, ch.keep = ch.out, None, None
                if req.error is not None:
                    continue  # the request was rejected while this chunk ran
                req.parts[ch.lo] = out
                req.chunks.remove((ch.lo, ch.hi))
                if not req.chunks:
                    outs = [req.parts[k] for k in sorted(req.parts)]
                    req.result = outs[0] if len(outs) == 1 else torch.cat(outs, 0)
                    req.parts = {}
                    req.t_done = self.now()
                    self.done.append(req)
            else:
                still.append(ch)
        self.inflight = still
        if self.queue:
            self._admit()  # also lets a chunk that just aged claim its reservation
        return retired

    # ----------------------------------------------------------- driver
    def serve(self, requests: typing.List[Request], timeout_s: float = 600.0):
        """Real-time loop: submit each request at its arrival time, poll until all are done."""
        pending = sorted(requests, key=lambda r: r.arrival)
        self.t0 = time.perf_counter()
        i = 0
        while i < len(pending) or self.queue or self.inflight:
            t = self.now()
            if t > timeout_s:
                raise TimeoutError("serving loop exceeded its time limit")
            while i < len(pending) and pending[i].arrival <= t:
                self.submit(pending[i])
                i += 1
            if not self.poll():
                nxt = pending[i].arrival - self.now() if i < len(pending) else 0.0
                time.sleep(min(max(nxt, 0.0), 2e-4) if nxt > 0 else 5e-5)
        torch.cuda.synchronize(self.device)
        self.arena = None
        return self.done


def latency_stats(reqs: typing.List[Request]):
    ok = [r for r in reqs if r.error is None]
    lat = np.array([r.latency for r in ok]) * 1e3
    svc = np.array([r.service for r in ok]) * 1e3

    def pct(a, q):
        return float(np.percentile(a, q)) if len(a) else None

    return {
        "requests": len(reqs),
        "rejected": len(reqs) - len(ok),
        "p50_ms": pct(lat, 50),
        "p99_ms": pct(lat, 99),
        "mean_ms": float(lat.mean()) if len(lat) else None,
        "service_p50_ms": pct(svc, 50),
        "service_p99_ms": pct(svc, 99),
    }


def poisson_requests(num: int, rate_hz: float, seed: int, device, n_lo=10, n_hi=28, n_mean=16,
                     b_max=128, d_lo=4, d_hi=16):
    """Seeded request stream: Poisson arrivals; n = n_lo + Geometric (mean n_mean,
    redrawn above n_hi); batch U{1..b_max}; depth U{d_lo..d_hi} of the
    RY/RZ + CNOT-ring ansatz with angles U[0, 2 pi)."""
    import qandle_amd as qandle
    from qandle_amd import workloads

    rng = np.random.default_rng(seed)
    gen = torch.Generator().manual_seed(seed)
    p = 1.0 / (n_mean - n_lo + 1)
    t = 0.0
    out = []
    for rid in range(num):
        t += rng.exponential(1.0 / rate_hz)
        while True:
            n = n_lo + int(rng.geometric(p)) - 1
            if n <= n_hi:
                break
        B = int(rng.integers(1, b_max + 1))
        depth = int(rng.integers(d_lo, d_hi + 1))
        w = workloads.hardware_efficient(n, B, depth, gen)
        circ = qandle.Circuit(w.layers, num_qubits=n).to(device)
        out.append(Request(rid, circ, {"x": w.inputs["x"].to(device)}, arrival=t))
    return out
