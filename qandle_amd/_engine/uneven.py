"""Batches of circuits with uneven widths and depths in one captured launch sequence.

Every circuit is lowered and planned on its own (its passes, tiles and words
are exactly those of a single-circuit run), then all circuits share:

* one coefficient table (all angles -> one ``qs_half_angle_coeffs``),
* one product-state factor table (``qs_factor_entries``),
* one state pool, one measurement-partial pool, one words pool,
* one work-item launch per *level*: level l runs pass l of every circuit that
  has one, as (descriptor, tile) items in a grid-stride loop; small circuits
  fit in one tile and finish in level 0,
* one batched MeasureProbability finish.

The backward is the same with adjoint descriptors: level l runs, for every
circuit, its reverse-sweep step l (its passes in reverse, then the gradient-only
passes of its embedding), the first step building the co-state from dL/dP.

The level sequence is captured into one hipGraph (torch.cuda.CUDAGraph on
ROCm records HIP graph nodes) and replayed: one graph launch per forward or
forward + backward, independent of the number of circuits.  Levels replace
the spin-waits a single persistent kernel would need between dependent tiles:
stream order guarantees pass l of every circuit is complete before level l+1
starts, and no workgroup ever waits on another.
"""
from __future__ import annotations

import dataclasses
import typing

import numpy as np
import torch

from . import fused
from . import lib as L
from . import planner as PL


@dataclasses.dataclass
class _Circ:
    prog: fused.Program
    n: int
    B: int
    d: int
    plan: PL.Plan
    gplan: typing.Optional[PL.Plan]
    ggates: list
    phi_off: int
    nphi: int
    x_off: int
    state_off: int          # complex64 elements
    out_off: int            # float elements of the (B, n) probabilities
    mpart_off: int          # float elements
    gpart_off: int          # float elements
    gout_off: int           # float elements of (rows, B) per-sample gradients
    gtot_off: int           # float elements of (rows,) totals
    efac_off: int           # float4 entries
    rows: int
    nrot: int


class UnevenBatch:
    """Forward (+ adjoint backward) of many independent circuits.

    ``circuits``: built ``qandle_amd.Circuit`` modules whose layers are an
    optional AngleEmbedding, rotations with their own (unnamed) angles, CNOTs
    and a final MeasureProbability.  ``inputs``: per circuit, the embedding
    input (B, d) or None.  Angles are snapshotted into ``self.phi`` (flat,
    remapped) at construction; ``set_angles`` refreshes them.
    """

    @classmethod
    def from_workloads(cls, workloads, device, **kw):
        """Batch straight from gate-list workloads (no operator modules are built)."""
        progs = [fused.program_from_spec(w.n, w.gates, w.embed) for w in workloads]
        return cls(progs, [w.inputs[w.embed[2]] for w in workloads], device, **kw)

    def __init__(self, circuits, inputs, device, R: int = 4, TB: int = 7, grid: int = 2048,
                 cost: PL.CostModel = None):
        self.device = torch.device(device)
        self.R, self.TB, self.grid = R, TB, grid
        self.cost = cost or fused.UNEVEN_COST
        lib = L.load()
        self.lib = lib
        self.circs: typing.List[_Circ] = []
        phi_vals, x_vals = [], []
        off = dict(phi=0, x=0, state=0, out=0, mpart=0, gpart=0, gout=0, gtot=0, efac=0)
        self.modules = circuits
        for ci, (circ, x) in enumerate(zip(circuits, inputs)):
            prog = circ if isinstance(circ, fused.Program) else fused.lower_circuit(circ.circuit)
            if prog is None or not prog.measure or prog.named_slots:
                raise ValueError(f"circuit {ci}: uneven batches take embedding + gates + MeasureProbability "
                                 "circuits without named rotation inputs")
            n = prog.n
            if n < R:
                raise ValueError(f"circuit {ci} has {n} qubits, fewer than the {R} register bits")
            if prog.init[0] == "embed":
                x = torch.as_tensor(x, dtype=torch.float32)
                if x.dim() == 1:
                    x = x.unsqueeze(0)
                B, d = x.shape[0], x.shape[1]
                if d != len(prog.init[2]):
                    raise ValueError(f"circuit {ci}: embedding expects {len(prog.init[2])} features")
            else:
                B, d = 1, 0
            pl = PL.plan(n, prog.gates, R, TB, self.cost)
            ggates, gplan = [], None
            if d:
                _, rot, bits, _ = prog.init
                ggates = [PL.Gate(PL.PGRAD + rot, b, -1, j) for j, b in enumerate(bits)]
                gplan = PL.plan(n, ggates, R, TB, self.cost)
            nrot = sum(1 for g in prog.gates if g.kind != PL.CNOT)
            rows = nrot + d
            mwb = lib.qs_fused_measure_work_bytes(B, n, R, TB) // 4
            gwb = lib.qs_adjoint_work_bytes(B, n, R, TB, rows) // 4
            prog.slot_index = {s_: i for i, s_ in enumerate(prog.param_slots)}
            c = _Circ(prog, n, B, d, pl, gplan, ggates, off["phi"], len(prog.param_slots), off["x"],
                      off["state"], off["out"], off["mpart"], off["gpart"], off["gout"], off["gtot"],
                      off["efac"], rows, nrot)
            self.circs.append(c)
            params = [prog.slots[i] for i in prog.param_slots]
            if params:
                phi_vals.append(torch.tensor(
                    [float(sl.module.remapping(sl.module.theta).detach().reshape(-1)[0]) if sl.kind == "param"
                     else sl.value for sl in params], dtype=torch.float32))
            if d:
                x_vals.append(x.reshape(-1))
            off["phi"] += c.nphi
            off["x"] += B * d
            off["state"] += B << n
            off["out"] += B * n
            off["mpart"] += mwb
            off["gpart"] += gwb
            off["gout"] += max(rows, 1) * B
            off["gtot"] += max(rows, 1)
            off["efac"] += B * n
        dev = self.device
        self.phi = (torch.cat(phi_vals) if phi_vals else torch.zeros(1)).to(dev)
        if self.phi.numel() == 0:
            self.phi = torch.zeros(1, device=dev)
        self.x = (torch.cat(x_vals) if x_vals else torch.zeros(1)).to(dev)
        self.cs = torch.empty(self.phi.numel(), 2, dtype=torch.float32, device=dev)
        self.efac = torch.empty(max(off["efac"], 1), 4, dtype=torch.float32, device=dev)
        self.state = torch.empty(max(off["state"], 1), dtype=torch.complex64, device=dev)
        self.out = torch.empty(max(off["out"], 1), dtype=torch.float32, device=dev)
        self.mpart = torch.empty(max(off["mpart"], 1), dtype=torch.float32, device=dev)
        self.totals = off
        self._build_factor_entries()
        self._build_forward()
        self._bw_built = False
        self.graphs = {}

    # ------------------------------------------------------------------ helpers
    @staticmethod
    def _addr(t: torch.Tensor, elem_off: int) -> int:
        return t.data_ptr() + elem_off * t.element_size()

    def _build_factor_entries(self):
        ent = []
        for c in self.circs:
            slot_of = {}
            rot = 0
            if c.d:
                _, rot, bits, _ = c.prog.init
                slot_of = {b: j for j, b in enumerate(bits)}
            for b in range(c.B):
                for pos in range(c.n):
                    j = slot_of.get(pos, -1)
                    ent.append((c.x_off + b * c.d + j) if j >= 0 else -1)
                    ent.append(rot)
        self.entries = torch.tensor(ent if ent else [-1, 0], dtype=torch.int32, device=self.device)
        self.nentries = len(ent) // 2

    def _upload_words(self, chunks):
        flat = np.concatenate(chunks) if chunks else np.zeros(1, np.int32)
        return torch.from_numpy(flat.astype(np.int32)).to(self.device)

    def _build_forward(self):
        lib, R, TB = self.lib, self.R, self.TB
        dbytes = lib.qs_pass_desc_bytes()
        mbytes = lib.qs_measure_desc_bytes()
        chunks, where = [], []
        pos = 0
        tables = PL.TableBuilder(1)
        for c in self.circs:
            for p in c.plan.passes:
                w = PL.encode_pass(p, c.prog.gates, R, TB, lambda s, c=c: c.phi_off + c.prog.slot_index[s],
                                      lambda s: False, tables)
                chunks.append(w)
                where.append(pos)
                pos += len(w)
        self.fw_words = self._upload_words(chunks)
        self.fw_chunks = chunks
        self.fw_tables = fused.upload_tables([], tables, self.device)
        self.fw_utab = torch.empty(max(tables.nu, 1) * 8, dtype=torch.float32, device=self.device)
        descs = bytearray()
        levels: typing.Dict[int, list] = {}
        k = 0
        ntiles_c = L.ctypes.c_int64(0)
        buf = (L.ctypes.c_char * dbytes)()
        for ci, c in enumerate(self.circs):
            st_ptr = self._addr(self.state, c.state_off)
            for pi, p in enumerate(c.plan.passes):
                last = pi == len(c.plan.passes) - 1
                lpos = np.asarray(p.L, dtype=np.int32)
                rc = lib.qs_pass_desc(
                    buf, L.ctypes.byref(ntiles_c), st_ptr, None if pi == 0 else st_ptr, c.B, c.n, R, TB,
                    lpos.ctypes.data_as(L.ctypes.POINTER(L.ctypes.c_int32)),
                    self._addr(self.fw_words, where[k]), len(p.subs),
                    L.ptr(self.cs), L.ptr(self.fw_utab), self._addr(self.efac, 4 * c.efac_off),
                    self._addr(self.mpart, c.mpart_off) if last else None, 1,
                )
                L.check(rc, f"qs_pass_desc(circuit {ci}, pass {pi})")
                levels.setdefault(pi, []).extend((k, t) for t in range(ntiles_c.value))
                descs += bytes(buf)
                k += 1
        self.fw_descs = torch.frombuffer(descs, dtype=torch.uint8).to(self.device)
        self.fw_levels = [torch.tensor(np.asarray(levels[l], dtype=np.int32).reshape(-1), device=self.device)
                          for l in sorted(levels)]
        self.fw_level_items = [len(levels[l]) for l in sorted(levels)]
        mdescs = bytearray()
        mbuf = (L.ctypes.c_char * mbytes)()
        for ci, c in enumerate(self.circs):
            lpos = np.asarray(c.plan.passes[-1].L, dtype=np.int32)
            L.check(lib.qs_measure_desc(mbuf, self._addr(self.out, c.out_off), self._addr(self.mpart, c.mpart_off),
                                        c.B, c.n, R, TB, lpos.ctypes.data_as(L.ctypes.POINTER(L.ctypes.c_int32))),
                    f"qs_measure_desc(circuit {ci})")
            mdescs += bytes(mbuf)
        self.m_descs = torch.frombuffer(mdescs, dtype=torch.uint8).to(self.device)
        self.max_n = max(c.n for c in self.circs)

    def _build_backward(self):
        lib, R, TB = self.lib, self.R, self.TB
        dev = self.device
        self.psi = torch.empty_like(self.state)
        self.lam = torch.empty_like(self.state)
        self.w = torch.zeros(max(self.totals["out"], 1), dtype=torch.float32, device=dev)
        self.gpart = torch.empty(max(self.totals["gpart"], 1), dtype=torch.float32, device=dev)
        self.gout = torch.empty(max(self.totals["gout"], 1), dtype=torch.float32, device=dev)
        self.gtot = torch.empty(max(self.totals["gtot"], 1), dtype=torch.float32, device=dev)
        chunks, steps = [], []
        pos = 0
        tables = PL.TableBuilder(1)
        for c in self.circs:
            rows = {}
            for gi, g in enumerate(c.prog.gates):
                if g.kind != PL.CNOT:
                    rows[gi] = len(rows)
            seq = []
            for p in reversed(c.plan.passes):
                w = PL.encode_pass(p, c.prog.gates, R, TB, lambda s, c=c: c.phi_off + c.prog.slot_index[s],
                                      lambda s: False, tables, reverse=True, row_of=rows.__getitem__)
                seq.append((p, w))
            if c.gplan is not None:
                for p in c.gplan.passes:
                    w = PL.encode_pass(p, c.ggates, R, TB, lambda s: 0, lambda s: False, tables,
                                          reverse=True, row_of=lambda gi, c=c: c.nrot + gi)
                    seq.append((p, w))
            st = []
            for p, w in seq:
                chunks.append(w)
                st.append((p, pos, len(p.subs)))
                pos += len(w)
            steps.append(st)
        self.bw_words = self._upload_words(chunks)
        self.bw_chunks = chunks
        self.bw_steps = steps
        self.bw_tables = fused.upload_tables([], tables, dev)
        self.bw_utab = torch.empty(max(tables.nu, 1) * 8, dtype=torch.float32, device=dev)
        self.bw_qtab = torch.empty(max(tables.nq, 1) * 8, dtype=torch.float32, device=dev)
        dbytes = lib.qs_adjoint_desc_bytes()
        buf = (L.ctypes.c_char * dbytes)()
        ntiles_c = L.ctypes.c_int64(0)
        descs = bytearray()
        levels: typing.Dict[int, list] = {}
        k = 0
        for ci, (c, st) in enumerate(zip(self.circs, steps)):
            for li, (p, wpos, nsub) in enumerate(st):
                lpos = np.asarray(p.L, dtype=np.int32)
                rc = lib.qs_adjoint_desc(
                    buf, L.ctypes.byref(ntiles_c), self._addr(self.psi, c.state_off), self._addr(self.lam, c.state_off),
                    c.B, c.n, R, TB, lpos.ctypes.data_as(L.ctypes.POINTER(L.ctypes.c_int32)),
                    self._addr(self.bw_words, wpos), nsub, L.ptr(self.cs), L.ptr(self.bw_utab),
                    L.ptr(self.bw_qtab), self._addr(self.gpart, c.gpart_off), c.rows,
                    self._addr(self.w, c.out_off) if li == 0 else None,
                )
                L.check(rc, f"qs_adjoint_desc(circuit {ci}, step {li})")
                levels.setdefault(li, []).extend((k, t) for t in range(ntiles_c.value))
                descs += bytes(buf)
                k += 1
        self.bw_descs = torch.frombuffer(descs, dtype=torch.uint8).to(dev)
        self.bw_levels = [torch.tensor(np.asarray(levels[l], dtype=np.int32).reshape(-1), device=dev)
                          for l in sorted(levels)]
        self.bw_level_items = [len(levels[l]) for l in sorted(levels)]
        fbytes = lib.qs_adjoint_finish_desc_bytes()
        fbuf = (L.ctypes.c_char * fbytes)()
        fdescs = bytearray()
        for ci, c in enumerate(self.circs):
            L.check(lib.qs_adjoint_finish_desc(fbuf, self._addr(self.gout, c.gout_off), self._addr(self.gtot, c.gtot_off),
                                               self._addr(self.gpart, c.gpart_off), c.rows, c.B, c.n, R, TB),
                    f"qs_adjoint_finish_desc(circuit {ci})")
            fdescs += bytes(fbuf)
        self.f_descs = torch.frombuffer(fdescs, dtype=torch.uint8).to(dev)
        # gradient scatter maps: parameter rows -> flat phi; embedding rows -> flat x
        pidx, prow = [], []
        xidx, xsrc = [], []
        for c in self.circs:
            rows = 0
            for gi, g in enumerate(c.prog.gates):
                if g.kind == PL.CNOT:
                    continue
                pidx.append(c.phi_off + c.prog.slot_index[g.slot])
                prow.append(c.gtot_off + rows)
                rows += 1
            for b in range(c.B):
                for j in range(c.d):
                    xidx.append(c.x_off + b * c.d + j)
                    xsrc.append(c.gout_off + (c.nrot + j) * c.B + b)
        self.p_index = torch.tensor(pidx or [0], dtype=torch.long, device=dev)
        self.p_rows = torch.tensor(prow or [0], dtype=torch.long, device=dev)
        self.np_rows = len(pidx)
        self.x_index = torch.tensor(xidx or [0], dtype=torch.long, device=dev)
        self.x_src = torch.tensor(xsrc or [0], dtype=torch.long, device=dev)
        self.nx = len(xidx)
        self.grad_phi = torch.zeros_like(self.phi)
        self.grad_x = torch.zeros_like(self.x)
        self._bw_built = True

    # ------------------------------------------------------------------ running
    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _compose(self, ws, utab, qtab, st):
        L.check(self.lib.qs_compose_tables(L.ptr(utab), None if qtab is None else L.ptr(qtab),
                                           L.ptr(ws.recipes), L.ptr(ws.offsets),
                                           ws.nrec, L.ptr(self.cs), st), "qs_compose_tables")

    def level_bytes(self):
        """Algorithmic HBM bytes of each forward level launch: a circuit's first pass
        writes its generated state (8 B per amplitude), every later pass reads and
        writes it (16 B); the measurement partials are negligible."""
        out = []
        for lvl in range(len(self.fw_level_items)):
            tot = 0
            for c in self.circs:
                if lvl < len(c.plan.passes):
                    tot += c.B * (2 ** c.n) * (8 if lvl == 0 else 16)
            out.append(tot)
        return out

    def run_forward(self, events=None):
        """Forward of every circuit.  ``events``: optional list; (start, end, bytes)
        HIP events on the launch stream are appended around every level launch."""
        lib, st = self.lib, self._stream()
        L.check(lib.qs_half_angle_coeffs(L.ptr(self.cs), L.ptr(self.phi), self.phi.numel(), st), "coeffs")
        L.check(lib.qs_factor_entries(L.ptr(self.efac), L.ptr(self.x), L.ptr(self.entries), self.nentries, st),
                "qs_factor_entries")
        self._compose(self.fw_tables, self.fw_utab, None, st)
        nbytes = self.level_bytes() if events is not None else None
        for lvl, (items, cnt) in enumerate(zip(self.fw_levels, self.fw_level_items)):
            if events is not None:
                ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                ev[0].record(torch.cuda.current_stream(self.device))
            L.check(lib.qs_items_pass(L.ptr(self.fw_descs), L.ptr(items), cnt, self.R, self.TB, self.grid, st),
                    "qs_items_pass")
            if events is not None:
                ev[1].record(torch.cuda.current_stream(self.device))
                events.append((ev[0], ev[1], nbytes[lvl]))
        L.check(lib.qs_items_measure_finish(L.ptr(self.m_descs), len(self.circs), self.max_n, st),
                "qs_items_measure_finish")

    def run_backward(self):
        """Adjoint sweep for the loss sum(w * P); ``self.w`` holds w (flat like ``self.out``)."""
        if not self._bw_built:
            self._build_backward()
        lib, st = self.lib, self._stream()
        self.psi.copy_(self.state)
        self._compose(self.bw_tables, self.bw_utab, self.bw_qtab, st)
        for items, cnt in zip(self.bw_levels, self.bw_level_items):
            L.check(lib.qs_items_adjoint(L.ptr(self.bw_descs), L.ptr(items), cnt, self.R, self.TB, self.grid, st),
                    "qs_items_adjoint")
        L.check(lib.qs_items_adjoint_finish(L.ptr(self.f_descs), len(self.circs), 8, st),
                "qs_items_adjoint_finish")
        self.grad_phi.zero_()
        if self.np_rows:
            self.grad_phi.index_add_(0, self.p_index, self.gtot[self.p_rows])
        if self.nx:
            self.grad_x.index_copy_(0, self.x_index, self.gout[self.x_src])

    def capture(self, backward: bool = True):
        """Record forward (and backward) into HIP graphs; replay with ``replay``."""
        if backward and not self._bw_built:
            self._build_backward()
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):  # warm-up outside capture
            self.run_forward()
            if backward:
                self.run_backward()
        torch.cuda.current_stream(self.device).wait_stream(s)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run_forward()
            if backward:
                self.run_backward()
        self.graphs["fb" if backward else "f"] = g
        return g

    def replay(self, backward: bool = True):
        self.graphs["fb" if backward else "f"].replay()

    # ------------------------------------------------------------------ results
    def probs(self, i: int) -> torch.Tensor:
        c = self.circs[i]
        return self.out[c.out_off: c.out_off + c.B * c.n].view(c.B, c.n)

    def set_loss_weights(self, ws):
        if not self._bw_built:
            self._build_backward()
        for c, w in zip(self.circs, ws):
            self.w[c.out_off: c.out_off + c.B * c.n].copy_(torch.as_tensor(w, dtype=torch.float32).reshape(-1))

    def param_grads(self, i: int) -> torch.Tensor:
        c = self.circs[i]
        return self.grad_phi[c.phi_off: c.phi_off + c.nphi]

    def input_grads(self, i: int) -> torch.Tensor:
        c = self.circs[i]
        return self.grad_x[c.x_off: c.x_off + c.B * c.d].view(c.B, c.d)

    @property
    def amplitude_updates(self) -> int:
        return sum((c.d + len(c.prog.gates)) * c.B * 2 ** c.n for c in self.circs)
