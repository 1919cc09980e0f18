"""Continuous-batching serving loop for circuit evaluations on one GPU.

The reference has no serving component; this is the request path a model
server puts in front of ``Circuit.forward``:

* Requests (a lowered circuit, its inputs, an arrival time) wait in a FIFO.
* Admission is by HBM footprint: a request is admitted when the resident
  bytes of everything in flight plus its own state (B x 2^n x 8 B, plus its
  measurement partials) stay under the cap and a stream slot is free.  A
  request whose batch alone exceeds the cap is split into batch chunks that
  each fit; a single sample that does not fit is rejected with an error.
* Admitted work is enqueued at once on one of several HIP streams (default
  priority) and overlaps on the device with everything else in flight; the
  host never blocks on a request: completion is detected by polling a HIP
  event, after which the request's memory is released and more requests are
  admitted (continuous batching, not a queue drained one batch at a time).
* Results are a pure function of (circuit, inputs): the kernel shape and
  fusion plan do not depend on load, and every reduction has a fixed order, so
  a request computes the same bits alone and in a mix.
* Requests run on the library's interpreting pass kernels: a new circuit is
  served at once instead of waiting for per-program kernels to compile.
* Latency = completion time - arrival time (host clock); p50/p99 reported with
  throughput.
"""
from __future__ import annotations

import collections
import dataclasses
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


@dataclasses.dataclass
class _Chunk:
    req: Request
    lo: int
    hi: int
    bytes: int
    event: typing.Any = None
    out: typing.Any = None
    keep: typing.Any = None


def state_footprint(n: int, B: int, R: int, TB: int) -> int:
    """HBM bytes a running evaluation of (n, B) holds: the state plus partials."""
    lib = L.load()
    return B * (2 ** n) * 8 + max(lib.qs_fused_measure_work_bytes(B, n, R, TB), 0) + B * n * 32 + 4096


class ServingEngine:
    def __init__(self, device, cap_bytes: float = 200e9, streams: int = 4, max_inflight: int = 64,
                 max_bypass_s: float = 2.0):
        self.device = torch.device(device)
        self.max_bypass_s = max_bypass_s
        self.cap = int(cap_bytes)
        self.streams = [torch.cuda.Stream(self.device) for _ in range(streams)]
        self.max_inflight = max_inflight
        self.queue: collections.deque = collections.deque()
        self.inflight: typing.List[_Chunk] = []
        self.resident = 0
        self.peak_resident = 0
        self.done: typing.List[Request] = []
        self.t0 = None
        self._rr = 0
        L.load()

    # ----------------------------------------------------------- admission
    def prepare(self, req: Request):
        """Lower + plan (once, before serving) and split over-cap batches."""
        prog = fused.lower_circuit(req.circuit.circuit)
        if prog is None or not prog.measure:
            raise ValueError("serving takes circuits that end in MeasureProbability")
        x = req.inputs[prog.init[3]] if prog.init[0] == "embed" else None
        req.n = prog.n
        req.B = 1 if x is None or x.dim() == 1 else x.shape[0]
        R, TB = fused.interpreted_shape(prog.n)
        fused.get_plan(prog, R, TB, fused.INTERPRETED_COST)
        per = state_footprint(req.n, 1, R, TB)
        req.footprint = state_footprint(req.n, req.B, R, TB)
        if per > self.cap:
            req.error = f"one sample needs {per} bytes, over the {self.cap}-byte admission cap"
            return req
        step = req.B if req.footprint <= self.cap else max(1, self.cap // per)
        req.chunks = [(lo, min(req.B, lo + step)) for lo in range(0, req.B, step)]
        return req

    def submit(self, req: Request):
        if req.error:
            req.t_admit = req.t_done = self.now()
            self.done.append(req)
            return
        req.t_queued = self.now()
        for lo, hi in req.chunks:
            R, TB = fused.interpreted_shape(req.n)
            self.queue.append(_Chunk(req, lo, hi, state_footprint(req.n, hi - lo, R, TB)))
        self._admit()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def _admit(self):
        """First fit in arrival order: a chunk that does not fit does not block later,
        smaller ones, unless it has waited longer than ``max_bypass_s`` (then nothing
        passes it, so large requests cannot starve)."""
        i = 0
        while i < len(self.queue) and len(self.inflight) < self.max_inflight:
            ch = self.queue[i]
            if self.resident + ch.bytes > self.cap and self.inflight:
                if i == 0 and ch.req.t_queued >= 0 and self.now() - ch.req.t_queued > self.max_bypass_s:
                    break  # the head has waited long enough: hold everything for it
                i += 1
                continue
            del self.queue[i]
            try:
                self._launch(ch)
            except torch.cuda.OutOfMemoryError:
                # the allocator could not place it after all: retry once memory frees up,
                # or reject it when nothing else is holding memory
                if self.inflight:
                    self.queue.insert(i, ch)
                    break
                self._reject(ch.req, "out of device memory at admission")
                i = 0  # the queue lost the request's other chunks: scan it again

    def _reject(self, req: Request, why: str):
        """Fail a request once: its queued chunks leave the queue, chunks of it still
        in flight are dropped when they retire (poll), and it is reported done now."""
        if req.error is not None:
            return
        req.error = why
        self.queue = collections.deque(c for c in self.queue if c.req is not req)
        req.chunks = []
        req.parts = {}
        req.t_done = self.now()
        self.done.append(req)

    def _launch(self, ch: _Chunk):
        req = ch.req
        s = self.streams[self._rr % len(self.streams)]
        self._rr += 1
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s), torch.no_grad():
            kw = dict(req.inputs)
            prog = fused.lower_circuit(req.circuit.circuit)
            if prog.init[0] == "embed":
                x = kw[prog.init[3]]
                kw[prog.init[3]] = x[ch.lo:ch.hi] if x.dim() == 2 else x
            inp = fused.gather_inputs(prog, req.circuit.circuit, None, kw)
            R, TB = fused.interpreted_shape(prog.n)
            cs = fused.build_coeffs(prog, inp)
            efac = fused.product_factors(prog, inp)
            # interpreting kernels: a request never waits for a program to compile
            _, probs = fused.execute(prog, inp, cs, efac, R, TB, cm=fused.INTERPRETED_COST, want_state=False,
                                     specialise=False)
            ch.out = probs
            ch.keep = (cs, efac)
            ev = torch.cuda.Event()
            ev.record(s)
            ch.event = ev
        if req.t_admit < 0:
            req.t_admit = self.now()
        self.resident += ch.bytes
        self.peak_resident = max(self.peak_resident, self.resident)
        self.inflight.append(ch)

    def poll(self) -> int:
        """Retire finished chunks, release their memory, admit more.  Returns #retired."""
        still, retired = [], 0
        for ch in self.inflight:
            if ch.event.query():
                retired += 1
                self.resident -= ch.bytes
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
        if retired:
            self._admit()
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
        return self.done


def latency_stats(reqs: typing.List[Request]):
    ok = [r for r in reqs if r.error is None]
    lat = np.array([r.latency for r in ok]) * 1e3
    return {
        "requests": len(reqs),
        "rejected": len(reqs) - len(ok),
        "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
        "mean_ms": float(lat.mean()) if len(lat) else None,
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
