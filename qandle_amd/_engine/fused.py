"""Circuit lowering and the fused-pass runner.

``lower_circuit`` turns a built ``UnsplittedCircuit`` (or ``SplittedCircuit``)
into a :class:`Program`: an initial state (|0..0>, the caller's state, or an
AngleEmbedding product state), a list of gates on bit positions with
coefficient slots, and whether MeasureProbability ends the circuit.  The
program is planned once per (register bits, thread bits, cost model) and
executed as a short sequence of ``qs_fused_pass`` launches.
"""
from __future__ import annotations

import dataclasses
import math
import typing
import weakref

import numpy as np
import torch

from . import lib as L
from . import planner as PL

DEFAULT_TB = 8
DEFAULT_R = 5


@dataclasses.dataclass
class Slot:
    """Where a rotation's angle comes from."""

    kind: str                     # "param" (remap(theta) of a module), "named", or "value" (a fixed angle)
    module: typing.Any = None
    name: str = ""
    value: float = 0.0
    sign: float = 1.0             # -1: the rotation is inverted (Invert of the module)


@dataclasses.dataclass
class Program:
    n: int
    gates: typing.List[PL.Gate]
    slots: typing.List[Slot]
    init: tuple                   # ("input",) | ("embed", rot, bits, name) | ("amplitude", name, normalize, pad)
    measure: bool
    post: str = None              # "abs2": MeasureJointProbability of the final state
    plans: dict = dataclasses.field(default_factory=dict)
    words_cache: dict = dataclasses.field(default_factory=dict)

    def __post_init__(self):
        self.param_slots = [i for i, s in enumerate(self.slots) if s.kind in ("param", "value")]
        self.named_slots = [i for i, s in enumerate(self.slots) if s.kind == "named"]


def _lower_into(layers, n, mapping, prog_gates, slots, state):
    """Append the lowering of ``layers`` (qubits mapped through ``mapping``)."""
    from .. import ansaetze, embeddings, measurements, operators, qcircuit

    for mod in layers:
        if isinstance(mod, (ansaetze.StronglyEntanglingLayerBuilt, ansaetze.StronglyEntanglingLayerBudget)):
            # plain gate sequences (src/qandle/ansaetze/stronglyentangling.py:109-110,
            # stronglyentangling_budget.py:89-90)
            if state.get("measured"):
                return False
            seq = mod.mods if isinstance(mod, ansaetze.StronglyEntanglingLayerBuilt) else mod.layers
            if not _lower_into(list(seq), n, mapping, prog_gates, slots, state):
                return False
            continue
        if isinstance(mod, operators.BuiltParametrizedOperator):
            q = mapping(mod.qubit)
            slot = Slot("named", name=mod.name) if mod.named else Slot("param", module=mod)
            slots.append(slot)
            prog_gates.append(PL.Gate(mod.KIND, n - 1 - q, -1, len(slots) - 1))
        elif isinstance(mod, operators.BuiltCNOT) or (
                isinstance(mod, operators.BuiltInvert) and isinstance(mod.target, operators.BuiltCNOT)):
            # CNOT is its own inverse
            g = mod if isinstance(mod, operators.BuiltCNOT) else mod.target
            prog_gates.append(PL.Gate(PL.CNOT, n - 1 - mapping(g.t), n - 1 - mapping(g.c)))
        elif isinstance(mod, operators.BuiltSWAP) or (
                isinstance(mod, operators.BuiltInvert) and isinstance(mod.target, operators.BuiltSWAP)):
            # SWAP(a, b) = CNOT(a, b) CNOT(b, a) CNOT(a, b), an exact permutation (its own inverse)
            g = mod if isinstance(mod, operators.BuiltSWAP) else mod.target
            if g.a != g.b:
                pa, pb = n - 1 - mapping(g.a), n - 1 - mapping(g.b)
                prog_gates += [PL.Gate(PL.CNOT, pb, pa), PL.Gate(PL.CNOT, pa, pb), PL.Gate(PL.CNOT, pb, pa)]
        elif isinstance(mod, operators.BuiltCZ) or (
                isinstance(mod, operators.BuiltInvert) and isinstance(mod.target, operators.BuiltCZ)):
            # CZ = H_t CNOT H_t with H = i RY(pi/2) RZ(pi) and the i^2 = -1 undone by
            # RZ(2 pi) = -I: fixed-angle rotations that fuse into the target's
            # per-bit unitaries, and a CNOT (its own inverse)
            g = mod if isinstance(mod, operators.BuiltCZ) else mod.target
            pt, pc = n - 1 - mapping(g.t), n - 1 - mapping(g.c)
            for kind, v in ((PL.RZ, math.pi), (PL.RY, math.pi / 2), (PL.CNOT, None), (PL.RZ, math.pi),
                            (PL.RY, math.pi / 2), (PL.RZ, 2 * math.pi)):
                if kind == PL.CNOT:
                    prog_gates.append(PL.Gate(PL.CNOT, pt, pc))
                else:
                    slots.append(Slot("value", value=v))
                    prog_gates.append(PL.Gate(kind, pt, -1, len(slots) - 1))
        elif (isinstance(mod, operators.BuiltInvert) and isinstance(mod.target, operators.BuiltParametrizedOperator)
              and not mod.target.named):
            # R_P(phi)^-1 = R_P(-phi) (the reference inverts target.to_matrix(), src/qandle/operators.py:445-447)
            t = mod.target
            slots.append(Slot("param", module=t, sign=-1.0))
            prog_gates.append(PL.Gate(t.KIND, n - 1 - mapping(t.qubit), -1, len(slots) - 1))
        elif isinstance(mod, embeddings.AngleEmbeddingBuilt):
            # the embedding discards the incoming state: everything before it is dead
            prog_gates.clear()
            slots.clear()
            bits = [n - 1 - mapping(q) for q in mod.qubits]
            state["init"] = ("embed", mod.kind, bits, mod.name)
        elif isinstance(mod, measurements.MeasureProbabilityBuilt):
            if mod.num_qubits != n or state.get("measured"):
                return False
            state["measured"] = True
        elif isinstance(mod, (measurements.MeasureAllAbsolute, measurements.MeasureState)):
            # terminal readouts of the final state (src/qandle/measurements.py:73-91)
            if state.get("measured") or state.get("post"):
                return False
            state["post"] = "abs2" if isinstance(mod, measurements.MeasureAllAbsolute) else "state"
        elif isinstance(mod, embeddings.AmplitudeEmbeddingBuilt):
            # the embedding replaces the state (src/qandle/embeddings.py:50-61): everything before it is dead;
            # its amplitudes fill the whole register, so the qubit map must be the identity
            if mod.num_qubits != n or any(mapping(q) != q for q in range(n)):
                return False
            prog_gates.clear()
            slots.clear()
            state["init"] = ("amplitude", mod.name, mod.normalize, mod.pad_with)
        elif isinstance(mod, measurements.MeasurementDummy):
            pass  # a decomposed measurement leaves the state as it is
        elif isinstance(mod, qcircuit.Circuit):
            if mod.num_qubits != n or state.get("measured"):
                return False
            inner = mod.circuit
            if isinstance(inner, qcircuit.SplittedCircuit):
                for sc in inner.subcircuits:
                    inv = sc.inverse
                    if not _lower_into(sc.layers, n, lambda s, inv=inv: mapping(inv[s]), prog_gates, slots, state):
                        return False
            elif not _lower_into(inner.layers, n, mapping, prog_gates, slots, state):
                return False
        else:
            return False
        if state.get("measured") and mod is not layers[-1] and not isinstance(mod, measurements.MeasureProbabilityBuilt):
            return False
        if state.get("post") and mod is not layers[-1] and not isinstance(
                mod, (measurements.MeasureAllAbsolute, measurements.MeasureState)):
            return False
    return True


# built circuit -> its Program.  The program caches plans, device tables and
# compiled-kernel handles (ctypes pointers), so it lives beside the module, not
# on it: the module itself stays deep-copyable and picklable like any
# nn.Module, and a copy lowers its own program (its slots then point at the
# copy's parameters).
_PROGRAMS: "weakref.WeakKeyDictionary" = weakref.WeakKeyDictionary()


def lower_circuit(circuit) -> typing.Optional[Program]:
    """Program for a built circuit, or None when a layer is outside the program format."""
    from .. import qcircuit

    if circuit in _PROGRAMS:
        return _PROGRAMS[circuit]
    n = circuit.num_qubits
    gates: typing.List[PL.Gate] = []
    slots: typing.List[Slot] = []
    state = {"init": ("input",)}
    ident = lambda q: q  # noqa: E731
    if isinstance(circuit, qcircuit.SplittedCircuit):
        ok = True
        for sc in circuit.subcircuits:
            inv = sc.inverse
            ok = ok and _lower_into(sc.layers, n, lambda s, inv=inv: inv[s], gates, slots, state)
    else:
        layers = list(circuit.layers)
        ok = _lower_into(layers, n, ident, gates, slots, state)
        # MeasureProbability must be the last layer
        from .. import measurements

        if ok and state.get("measured"):
            ok = isinstance(layers[-1], measurements.MeasureProbabilityBuilt) or isinstance(layers[-1], qcircuit.Circuit)
        if ok and state.get("post"):
            ok = isinstance(layers[-1], (measurements.MeasureAllAbsolute, measurements.MeasureState))
    if not ok or n > 40 or n < 1:
        _PROGRAMS[circuit] = None
        return None
    post = state.get("post")
    prog = Program(n, gates, slots, state["init"], bool(state.get("measured")), "abs2" if post == "abs2" else None)
    _PROGRAMS[circuit] = prog
    return prog


def program_from_spec(n: int, gates, embed=None, measure: bool = True) -> Program:
    """Program straight from a gate list, without building operator modules.

    ``gates``: (kind, qubit, angle) with kind in rx/ry/rz, or ("cnot", (c, t));
    ``embed``: (rot in "xyz", qubits, input name) or None (start from |0..0>).
    Angles are fixed values (no remapping)."""
    kinds = {"rx": PL.RX, "ry": PL.RY, "rz": PL.RZ}
    rot_kind = {"x": 0, "y": 1, "z": 2}
    pg, slots = [], []
    for g in gates:
        if g[0] == "cnot":
            c, t = g[1]
            pg.append(PL.Gate(PL.CNOT, n - 1 - t, n - 1 - c))
        else:
            slots.append(Slot("value", value=float(g[2])))
            pg.append(PL.Gate(kinds[g[0]], n - 1 - g[1], -1, len(slots) - 1))
    init = ("input",)
    if embed is not None:
        rot, qubits, name = embed
        init = ("embed", rot_kind[rot], [n - 1 - q for q in qubits], name)
    return Program(n, pg, slots, init, measure)


def readout_fold(prog: Program):
    """(number of gates kept, readout masks) when CNOTs end a measured program,
    else None.

    A CNOT maps basis state x to x with its target bit XOR-ed by its control
    bit, so a network of CNOTs is an invertible linear map over GF(2) and the
    probabilities after it are those before it, permuted.  MeasureProbability
    of the bit at position p after the network is therefore P(parity(a_p & x)
    = 0) of the state x before it, with a_p (a mask over bit positions) read
    off the network: starting from a_p = 1 << p, each CNOT from the last to
    the first adds its control to every mask that holds its target.  The
    trailing CNOTs of a circuit (the entangling layer that ends
    src/qandle/operators.py:557-558 ladders) then cost no pass at all."""
    if not prog.measure:
        return None
    hit = prog.words_cache.get("readout_fold", False)
    if hit is not False:
        return hit
    m = len(prog.gates)
    while m > 0 and prog.gates[m - 1].kind == PL.CNOT:
        m -= 1
    hit = None
    if 0 < m < len(prog.gates):
        masks = [1 << p for p in range(prog.n)]
        for g in reversed(prog.gates[m:]):
            for p in range(prog.n):
                if (masks[p] >> g.pos) & 1:
                    masks[p] ^= 1 << g.ctrl
        hit = (m, masks)
    prog.words_cache["readout_fold"] = hit
    return hit


def get_plan(prog: Program, R: int, TB: int, cm: PL.CostModel = None, fold: bool = False) -> PL.Plan:
    """Plan of the program's gates, or with ``fold`` of the gates that
    readout_fold keeps (the trailing CNOTs then live in the readout masks)."""
    cm = cm or DEFAULT_COST
    rf = readout_fold(prog) if fold else None
    key = (R, TB, None if cm is None else dataclasses.astuple(cm), rf is not None)
    pl = prog.plans.get(key)
    if pl is None:
        gates = prog.gates if rf is None else prog.gates[:rf[0]]
        pl = PL.plan(prog.n, gates, R, TB, cm)
        prog.plans[key] = pl
    return pl


def folds(prog: Program, specialise: bool) -> bool:
    """Whether a run folds the program's trailing CNOTs into its readout (the
    run-time specialised kernels only)."""
    return bool(specialise) and readout_fold(prog) is not None


# fusion budgets: the specialised kernels make per-gate work cheap, so passes
# are packed to the planner's structural limit; the interpreting kernels pay
# a dispatch per op and are kept lighter
DEFAULT_COST = PL.CostModel(budget=1024.0, min_low=4, lookahead=True)
# forward + adjoint: the lookahead's deeper passes (up to 14 sub-passes) make adjoint steps of 4 ms
# and measured no faster (20.5 vs 17-21 ms at 16 qubits x 256 x 20 layers), so the greedy schedule
GRAD_COST = PL.CostModel(budget=1024.0, min_low=4)
INTERPRETED_COST = PL.CostModel(budget=256.0, min_low=4)
# uneven batches (interpreting item kernels, one launch per level): passes packed to the structural
# limit, so fewer levels: 4.41 vs 6.25 ms per forward of the 2048-circuit batch (budget 160), 24.1 vs
# 30.8 ms forward + backward; 5 levels instead of 13
UNEVEN_COST = PL.CostModel(budget=1024.0, min_low=4)


def choose_config(prog: Program, inp, grad: bool = False):
    """(R, TB, cost model) for a call: the autotuned choice when enabled, else
    the defaults for a forward-only call or, with ``grad``, for forward + adjoint."""
    from .. import config

    if getattr(config, "AUTOTUNE", False):
        from . import autotune

        hit = autotune.lookup(prog, inp.B, inp.device)
        if hit is None:
            with torch.no_grad():
                hit, _ = autotune.tune(prog, inp)
        return hit
    R, TB = choose_shape(prog.n, grad)
    return R, TB, (GRAD_COST if grad else DEFAULT_COST)


def choose_shape(n: int, grad: bool = False):
    """(R, TB) of the specialised kernels for a circuit width.

    Forward only: from 13 qubits, tiles of 2^13 amplitudes over 256 threads
    with 32 amplitudes per thread (the fewest passes).  With gradients: from
    12 qubits, 2^12-amplitude tiles with 16 amplitudes per thread (the adjoint
    holds psi and lambda, and this shape keeps two waves per SIMD).  From 5
    qubits, 2^11-amplitude tiles of one wave; below that, every bit in
    registers of a 256-thread tile holding several samples."""
    if n >= 13 and not grad:
        return 5, 8
    if n >= 12:
        return 4, 8
    if n >= 5:
        return 5, 6
    return n, DEFAULT_TB


def interpreted_shape(n: int):
    """(R, TB) of the interpreting kernels (serving): one-wave 2^11-amplitude tiles."""
    return (5, 6) if n >= 5 else (n, DEFAULT_TB)


# ----------------------------------------------------------------------------
# runtime
# ----------------------------------------------------------------------------

@dataclasses.dataclass
class Inputs:
    B: int
    batched: bool
    device: torch.device
    phi_param: torch.Tensor        # (S,) remapped angles of unnamed gates (autograd-tracked)
    named: typing.Dict[str, torch.Tensor]
    embed_x: typing.Optional[torch.Tensor]
    state: typing.Optional[torch.Tensor]


def gather_inputs(prog: Program, circuit, state, kwargs) -> Inputs:
    device = None
    B, batched = 1, False
    if state is not None:
        L.require_device(state, "state")
        device = state.device
        if state.dim() == 2:
            B, batched = state.shape[0], True
        elif state.dim() != 1:
            raise RuntimeError("state must be (2^n,) or (B, 2^n)")
        if state.shape[-1] != 2 ** prog.n:
            raise RuntimeError(f"state has {state.shape[-1]} amplitudes, circuit expects {2 ** prog.n}")
    named = {}
    for i in prog.named_slots:
        name = prog.slots[i].name
        if name not in named:
            v = kwargs[name]
            if not isinstance(v, torch.Tensor):
                v = torch.tensor(v, dtype=torch.float32)
            named[name] = v
    embed_x = None
    if prog.init[0] == "amplitude":
        # the embedded amplitudes become the input state (autograd-tracked through the embedding)
        from .plain import AmplitudeEmbedFn

        _, name, normalize, pad = prog.init
        xa = kwargs[name]
        L.require_device(xa, f"input '{name}'")
        N = 2 ** prog.n
        if pad is None and xa.shape[-1] != N:
            raise RuntimeError(f"AmplitudeEmbedding without pad_with needs {N} features, got {xa.shape[-1]}")
        if xa.dim() not in (1, 2):
            raise RuntimeError("AmplitudeEmbedding input must be (d,) or (B, d)")
        batched_a = xa.dim() == 2
        x2 = xa.reshape(-1, xa.shape[-1]).to(torch.float32)
        state = AmplitudeEmbedFn.apply(x2, prog.n, 0.0 if pad is None else float(pad), bool(normalize))
        device = state.device
        B, batched = state.shape[0], batched_a
    if prog.init[0] == "embed":
        embed_x = kwargs[prog.init[3]]
        L.require_device(embed_x, f"input '{prog.init[3]}'")
        device = embed_x.device
        if embed_x.shape[-1] != len(prog.init[2]):
            raise RuntimeError(f"AngleEmbedding expects {len(prog.init[2])} features, got {embed_x.shape[-1]}")
        # the embedding discards the incoming state, so only its input sets the batch
        if embed_x.dim() == 2:
            B, batched = embed_x.shape[0], True
        elif embed_x.dim() == 1:
            B, batched = 1, False
        else:
            raise RuntimeError("AngleEmbedding input must be (d,) or (B, d)")
    if device is None:
        for v in named.values():
            if v.device.type == "cuda":
                device = v.device
                break
    if device is None:
        for p in circuit.parameters():
            device = p.device
            break
    if device is None or device.type != "cuda":
        raise L.EngineError(
            "qandle_amd runs on the MI355X HIP engine only; pass a state or inputs on a HIP device "
            "or move the circuit to one (module.to('cuda'))."
        )
    # named (B,) inputs batch the circuit; a batch of one broadcasts against them
    # (the reference's bmm broadcasting, src/qandle/operators.py:294-297)
    for name in list(named):
        v = named[name].to(device=device, dtype=torch.float32)
        if v.dim() > 1:
            raise RuntimeError(f"named input '{name}' must be 0-d or (B,)")
        if v.dim() == 1:
            if B == 1:
                B = v.shape[0]
            elif v.shape[0] == 1:
                v = v.expand(B)
            elif v.shape[0] != B:
                raise RuntimeError(f"named input '{name}' has {v.shape[0]} entries, batch is {B}")
            batched = True
        named[name] = v
    for name, v in named.items():
        if v.dim() == 1 and v.shape[0] != B:
            if v.shape[0] == 1:
                named[name] = v.expand(B)
            else:
                raise RuntimeError(f"named input '{name}' has {v.shape[0]} entries, batch is {B}")
    return Inputs(B, batched, device, param_angles(prog, device), named, embed_x, state)


def _first_angle(theta: torch.Tensor) -> torch.Tensor:
    return theta.reshape(1) if theta.numel() == 1 else theta.reshape(-1)[:1]


def _gather_thetas(thetas, cache=None) -> torch.Tensor:
    """One (len,) tensor of the rotations' angles (the first element of each theta),
    in order, with as few autograd nodes as the shapes allow.

    A circuit holds one theta Parameter per gate (the reference's state_dict
    layout), so a deep circuit has hundreds of them; one reshape per theta costs
    more host time than the circuit's kernels.  0-d thetas (angles given as
    numbers, ansatz parameters) go through one ``stack``, one-element 1-d thetas
    (default ``torch.rand(1)`` draws) through one ``cat``, anything else through a
    reshape; mixed shapes are put back in order with one gather."""
    zero = [i for i, t in enumerate(thetas) if t.dim() == 0]
    if len(zero) == len(thetas):
        return torch.stack(thetas)
    one = [i for i, t in enumerate(thetas) if t.dim() == 1 and t.shape[0] == 1]
    if len(one) == len(thetas):
        return torch.cat(thetas)
    rest = [i for i in range(len(thetas)) if i not in set(zero) | set(one)]
    parts, order = [], []
    if zero:
        parts.append(torch.stack([thetas[i] for i in zero]))
        order += zero
    if one:
        parts.append(torch.cat([thetas[i] for i in one]))
        order += one
    if rest:
        parts.append(torch.cat([_first_angle(thetas[i]) for i in rest]))
        order += rest
    flat = torch.cat(parts)
    key = ("theta_order", str(flat.device), tuple(order))
    idx = None if cache is None else cache.get(key)
    if idx is None:
        pos = [0] * len(order)
        for j, i in enumerate(order):
            pos[i] = j
        idx = L.upload(pos, torch.long, flat.device)
        if cache is not None:
            cache[key] = idx
    return flat.index_select(0, idx)


def param_angles(prog: Program, device) -> torch.Tensor:
    """Remapped angles of the program's unnamed rotations, in slot order.

    Each gate's angle is remapping(theta) as in the reference
    (src/qandle/operators.py:265-275).  The qw_map remappings are elementwise,
    so the thetas that share one are stacked and remapped in one launch (and
    their gradients come back through one stack) instead of one launch per
    gate; any other remapping callable is applied per gate.  Fixed angles are
    uploaded once per program and device."""
    from .. import qw_map

    params = [prog.slots[i] for i in prog.param_slots]
    if not params:
        return torch.zeros(0, dtype=torch.float32, device=device)
    elementwise = (qw_map.none, qw_map.tanh)
    pieces, order = [], []
    groups: typing.Dict[typing.Any, list] = {}
    values = []
    for k, sl in enumerate(params):
        if sl.kind == "param" and sl.module.remapping in elementwise:
            groups.setdefault(sl.module.remapping, []).append(k)
        elif sl.kind == "param":
            pieces.append(sl.module.remapping(sl.module.theta).reshape(-1)[:1].to(device=device, dtype=torch.float32))
            order.append([k])
        else:
            values.append(k)
    for fn, ks in groups.items():
        # one stack/cat of the thetas: their gradients come back as views of one buffer,
        # with no per-parameter kernel
        th = _gather_thetas([params[k].module.theta for k in ks], prog.words_cache)
        pieces.append(fn(th).to(device=device, dtype=torch.float32))
        order.append(ks)
    if values:
        key = ("values", str(device))
        vt = prog.words_cache.get(key)
        if vt is None:
            vt = L.upload([params[k].value for k in values], torch.float32, device)
            prog.words_cache[key] = vt
        pieces.append(vt)
        order.append(values)
    flat = [k for ks in order for k in ks]
    phi = pieces[0] if len(pieces) == 1 else torch.cat(pieces)
    if flat != list(range(len(params))):
        key = ("perm", str(device), tuple(flat))
        inv = prog.words_cache.get(key)
        if inv is None:
            pos = [0] * len(flat)
            for i, k in enumerate(flat):
                pos[k] = i
            inv = L.upload(pos, torch.long, device)
            prog.words_cache[key] = inv
        phi = phi.index_select(0, inv)
    if any(sl.sign != 1.0 for sl in params):
        key = ("signs", str(device))
        sg = prog.words_cache.get(key)
        if sg is None:
            sg = L.upload([sl.sign for sl in params], torch.float32, device)
            prog.words_cache[key] = sg
        phi = phi * sg
    return phi


def coefficient_layout(prog: Program, B: int):
    """float2 index of every slot in the coefficient table, and its per-sample flag."""
    index = {}
    per_sample = {}
    off = 0
    for i in prog.param_slots:
        index[i] = off
        per_sample[i] = False
        off += 1
    name_off = {}
    for i in prog.named_slots:
        nm = prog.slots[i].name
        if nm not in name_off:
            name_off[nm] = off
            off += B
        index[i] = name_off[nm]
        per_sample[i] = True
    return index, per_sample, off, name_off


def rotation_rows(prog: Program):
    """Gradient row of every rotation gate (program order)."""
    rows = {}
    for gi, g in enumerate(prog.gates):
        if g.kind != PL.CNOT:
            rows[gi] = len(rows)
    return rows


@dataclasses.dataclass
class WordSet:
    """Encoded passes plus the recipes of the fused-unitary tables they index."""

    passes: list                  # [(int32 words on the device, sub-pass count, lpos int32 numpy)]
    recipes: torch.Tensor         # int32 recipe words (qs_compose_tables)
    offsets: torch.Tensor         # int32 start of every recipe
    nrec: int
    nu: int                       # U entries
    nq: int                       # gradient-generator entries
    host: list = None             # the same passes with their words as host numpy arrays

    def host_passes(self):
        """The passes with host words (what the kernel compilers read; no device copy back)."""
        return self.host if self.host is not None else self.passes


def upload_tables(passes, tables: PL.TableBuilder, device, host=None) -> WordSet:
    rec = L.upload(tables.recipes(), torch.int32, device)
    off = L.upload(np.asarray(tables.offsets or [0], dtype=np.int32), torch.int32, device)
    return WordSet(passes, rec, off, tables.count, tables.nu, tables.nq, host)


def compose_tables(ws: WordSet, cs: torch.Tensor, generators: bool = True):
    """(utab, qtab): the fused 2x2 unitaries and (for the adjoint) their gradient
    generators for this call's angles; qtab is None without ``generators``."""
    utab = torch.empty(max(ws.nu, 1) * 8, dtype=torch.float32, device=cs.device)
    qtab = torch.empty(max(ws.nq, 1) * 8, dtype=torch.float32, device=cs.device) if generators else None
    L.check(L.load().qs_compose_tables(L.ptr(utab), None if qtab is None else L.ptr(qtab), L.ptr(ws.recipes),
                                       L.ptr(ws.offsets), ws.nrec, L.ptr(cs), L.stream_of(cs)),
            "qs_compose_tables")
    return utab, qtab


def _encode(p, gates, pl, index, per_sample, tables, rows=None):
    w = PL.encode_pass(p, gates, pl.R, pl.TB, index, per_sample, tables,
                       reverse=rows is not None, row_of=rows, keep_waves=pl.n >= pl.R + pl.TB)
    return w, len(p.subs), np.asarray(p.L, dtype=np.int32)


def pass_words(prog: Program, pl: PL.Plan, B: int, device) -> WordSet:
    """Forward words of every pass (cached per plan, batch and device)."""
    key = (id(pl), B, str(device), "fwd")
    cached = prog.words_cache.get(key)
    if cached is not None:
        return cached
    index, per_sample, _, _ = coefficient_layout(prog, B)
    tables = PL.TableBuilder(B)
    out, host = [], []
    for p in pl.passes:
        w, nsub, lpos = _encode(p, prog.gates, pl, index.__getitem__, per_sample.__getitem__, tables)
        out.append((L.upload(w, torch.int32, device), nsub, lpos))
        host.append((w, nsub, lpos))
    ws = upload_tables(out, tables, device, host)
    prog.words_cache[key] = ws
    return ws


def embed_grad_gates(prog: Program):
    """Gradient-only probes, one per embedded bit (Pauli of the embedding rotation)."""
    _, rot, bits, _ = prog.init
    return [PL.Gate(PL.PGRAD + rot, b, -1, j) for j, b in enumerate(bits)]


def adjoint_words(prog: Program, pl: PL.Plan, B: int, device, embed: bool, cm: PL.CostModel = None):
    """Reverse-sweep words: the passes in reverse (rotation rows in program
    order), then, with ``embed``, the probe passes of the embedded bits (rows
    after the rotations).  Returns (WordSet, number of reversed forward passes)."""
    key = (id(pl), B, str(device), "adj", embed)
    cached = prog.words_cache.get(key)
    if cached is not None:
        return cached
    index, per_sample, _, _ = coefficient_layout(prog, B)
    rows = rotation_rows(prog)
    tables = PL.TableBuilder(B)
    out, host = [], []
    for p in reversed(pl.passes):
        w, nsub, lpos = _encode(p, prog.gates, pl, index.__getitem__, per_sample.__getitem__, tables,
                                rows.__getitem__)
        out.append((L.upload(w, torch.int32, device), nsub, lpos))
        host.append((w, nsub, lpos))
    nmain = len(out)
    if embed:
        gates = embed_grad_gates(prog)
        gplan = PL.plan(prog.n, gates, pl.R, pl.TB, cm or DEFAULT_COST)
        row0 = len(rows)
        for p in gplan.passes:
            w, nsub, lpos = _encode(p, gates, gplan, lambda s_: 0, lambda s_: False, tables,
                                    lambda gi: row0 + gi)
            out.append((L.upload(w, torch.int32, device), nsub, lpos))
            host.append((w, nsub, lpos))
    hit = (upload_tables(out, tables, device, host), nmain)
    prog.words_cache[key] = hit
    return hit


def build_coeffs(prog: Program, inp: Inputs) -> torch.Tensor:
    """(entries, 2) float32 (cos, sin) of every slot's half-angle, on the device."""
    from .plain import half_angle_coeffs

    parts = [inp.phi_param.detach()]
    _, _, total, name_off = coefficient_layout(prog, inp.B)
    for nm in name_off:
        v = inp.named[nm].detach()
        parts.append(v.expand(inp.B) if v.dim() == 0 else v)
    phi = torch.cat([p.reshape(-1) for p in parts]) if parts else torch.zeros(0, device=inp.device)
    if phi.numel() == 0:
        phi = torch.zeros(1, device=inp.device)
    return half_angle_coeffs(phi)


def product_factors(prog: Program, inp: Inputs) -> torch.Tensor:
    n = prog.n
    efac = torch.empty(inp.B, n, 2, 2, dtype=torch.float32, device=inp.device)
    lib = L.load()
    if prog.init[0] == "embed":
        _, rot, bits, _ = prog.init
        x = inp.embed_x.detach().to(torch.float32)
        if x.dim() == 1:
            x = x.unsqueeze(0)
        if x.shape[0] != inp.B:
            x = x.expand(inp.B, -1)
        x = x.contiguous()
        L.check(
            lib.qs_product_factors(L.ptr(efac), L.ptr(x), inp.B, n, L.i32_array(bits), len(bits), rot,
                                   x.shape[1], L.stream_of(efac)),
            "qs_product_factors",
        )
    else:
        L.check(lib.qs_product_factors(L.ptr(efac), None, inp.B, n, None, 0, 0, 0, L.stream_of(efac)),
                "qs_product_factors")
    return efac


def leading_rotations(prog: Program):
    """(program of the remaining gates, rotations by bit position) or None.

    A rotation on a qubit that no CNOT has touched yet acts on an unentangled
    qubit of a product state, so it can be applied to that qubit's 2-vector
    factor (qs_factor_apply) instead of to all 2^n amplitudes: the first
    layer of an embedding circuit then costs B * n small products.  The
    remaining program keeps the slots (and so the coefficient layout) of the
    full one.  None when nothing is absorbed or nothing would remain."""
    hit = prog.words_cache.get("leading", False)
    if hit is not False:
        return hit
    hit = None
    if prog.init[0] in ("embed", "input"):
        touched, keep, per_pos = set(), [], {}
        for g in prog.gates:
            if g.kind in (PL.RX, PL.RY, PL.RZ) and g.pos not in touched:
                per_pos.setdefault(g.pos, []).append(g)
                continue
            touched.update(g.positions)
            keep.append(g)
        if per_pos and keep:
            hit = (Program(prog.n, keep, prog.slots, prog.init, prog.measure, prog.post), per_pos)
    prog.words_cache["leading"] = hit
    return hit


def _leading_tables(prog: Program, per_pos, B: int, device):
    """Recipes of the leading rotations' per-position unitaries (absorb_factors),
    cached per batch and device."""
    key = ("leading_tables", B, str(device))
    hit = prog.words_cache.get(key)
    if hit is None:
        index, per_sample, _, _ = coefficient_layout(prog, B)
        tables = PL.TableBuilder(B)
        unit = np.full(prog.n, -1, dtype=np.int32)
        span = np.ones(prog.n, dtype=np.int32)
        for pos in sorted(per_pos):
            u, ps, _ = tables.add([(g.kind, index[g.slot], per_sample[g.slot]) for g in per_pos[pos]])
            unit[pos], span[pos] = u, (B if ps else 1)
        recipes = L.upload(tables.recipes(), torch.int32, device)
        offsets = L.upload(tables.offsets, torch.int32, device)
        hit = (recipes, offsets, tables.count, tables.nu, unit, span)
        prog.words_cache[key] = hit
    return hit


def absorb_factors(prog: Program, per_pos, efac: torch.Tensor, cs: torch.Tensor, B: int) -> torch.Tensor:
    """Product-state factors with the leading rotations applied (leading_rotations):
    their unitaries are composed per bit position on the device (one recipe
    per position, per sample when an input angle is among them) and applied by
    qs_factor_apply."""
    recipes, offsets, nrec, nu, unit, span = _leading_tables(prog, per_pos, B, efac.device)
    lib = L.load()
    st = L.stream_of(efac)
    utab = torch.empty(max(nu, 1) * 8, dtype=torch.float32, device=efac.device)
    L.check(lib.qs_compose_tables(L.ptr(utab), None, L.ptr(recipes), L.ptr(offsets), nrec, L.ptr(cs), st),
            "qs_compose_tables")
    out = torch.empty_like(efac)
    L.check(lib.qs_factor_apply(L.ptr(out), L.ptr(efac), L.ptr(utab), nu, L.i32_array(unit), L.i32_array(span),
                                B, prog.n, st), "qs_factor_apply")
    return out


def tiles_of(B: int, n: int, K: int) -> int:
    """Tiles of one pass (as qs_fused_pass counts them)."""
    if n >= K:
        return B << (n - K)
    spt = 1 << (K - n)
    return (B + spt - 1) // spt


def use_jit(n: int, R: int, TB: int) -> bool:
    from .. import config

    return bool(getattr(config, "SPECIALISE", False))


def specialised_kernels(prog: Program, pl: PL.Plan, ws: WordSet, B: int, gen_first: bool, fold: bool = False,
                        load: bool = True):
    """Run-time compiled kernels of this plan's passes (cached on the program);
    with ``fold`` the last pass writes readout partials (readout_fold).
    ``load`` False only compiles them into the code-object cache."""
    from .. import config
    from . import jit

    opts = getattr(config, "SPECIALISE_OPTIONS", {}) or {}
    key = (id(pl), B, gen_first, prog.measure, fold, tuple(sorted(opts.items())))
    hit = prog.words_cache.get(("jit",) + key)
    if hit is None:
        masks = readout_fold(prog)[1] if fold else None
        hit = jit.ProgramKernels(ws.host_passes(), prog.n, pl.R, pl.TB, gen_first, prog.measure, readout=masks,
                                 load=load, **opts)
        if load:
            prog.words_cache[("jit",) + key] = hit
    return hit


def forward_kernels(prog: Program, B: int, R: int, TB: int, cm: PL.CostModel = None, device="cpu",
                    load: bool = True):
    """The compiled kernels a measured forward without gradients from a product
    state runs (the program ``execute`` ends up with: leading rotations absorbed,
    trailing CNOTs folded into the readout), compiled now if the code-object cache
    does not hold them.  ``device`` only places the encoded words; the kernel
    sources depend on the program's structure, not on its angles, batch or device,
    so a worker process can build them ahead of a server that runs them later
    (``load`` False: compile into the code-object cache without loading them
    onto a device)."""
    from .. import config

    if getattr(config, "ABSORB_LEADING_ROTATIONS", True) and prog.init[0] in ("embed", "input"):
        lead = leading_rotations(prog)
        if lead is not None:
            prog = lead[0]
    fold = folds(prog, True)
    pl = get_plan(prog, R, TB, cm, fold)
    ws = pass_words(prog, pl, B, device)
    return specialised_kernels(prog, pl, ws, B, True, fold, load=load)


def warm(prog: Program, B: int, R: int, TB: int, cm: PL.CostModel = None, specialise: bool = None,
         device="cuda", want_state: bool = False, input_state: bool = False):
    """Builds, uploads and compiles what execute needs for this program and
    batch (the leading rotations' tables, the plan, the pass words and
    fused-unitary recipes, the compiled kernels and their lane-table index)
    without launching any work.  A caller that will launch on a busy stream
    (a serving loop) warms ahead, so that execute then does no host work
    that waits for the device.  ``input_state``: the call passes an initial
    state (no product-state factors)."""
    from .. import config

    n = prog.n
    jit_on = use_jit(n, R, TB) if specialise is None else specialise
    product = not (prog.init[0] in ("input", "amplitude") and input_state)
    if product and getattr(config, "ABSORB_LEADING_ROTATIONS", True):
        lead = leading_rotations(prog)
        if lead is not None:
            _leading_tables(prog, lead[1], B, device)
            prog = lead[0]
    fold = folds(prog, jit_on) and not want_state
    pl = get_plan(prog, R, TB, cm, fold)
    ws = pass_words(prog, pl, B, device)
    if jit_on:
        kern = specialised_kernels(prog, pl, ws, B, product, fold)
        if kern.lanes == "lds":
            kern.lane_index_on(device)


def execute(prog: Program, inp: Inputs, cs: torch.Tensor, efac, R: int, TB: int,
            cm: PL.CostModel = None, want_state: bool = True, checkpoints=None, events=None,
            specialise: bool = None, fold: bool = None, out: torch.Tensor = None):
    """Runs every pass.  Returns (state or None, probs or None).

    ``checkpoints(i, state)``: called after pass i with the state it wrote; when
    it returns a tensor (same shape), the passes continue in that one and the
    caller keeps ``state`` as it is (a checkpoint without a copy).
    ``out``: the (B, 2^n) complex64 device tensor the passes run in (the
    returned state); None allocates it.  A caller that launches on a side
    stream can take it from the pool of the stream it allocates on (and
    record it on the side stream), so one pool serves every stream.

    ``events``: optional list; a (start, end) pair of HIP events recorded on the
    launch stream around every pass kernel is appended to it.  ``specialise``:
    run compiled-per-program kernels (None: config.SPECIALISE) or the library's
    interpreting kernels (False; no compilation on first use of a program).
    ``fold``: fold trailing CNOTs into the readout of a measured program
    (readout_fold; compiled kernels only), so that the state left behind is
    the one before those CNOTs; None folds when the state is not wanted."""
    from .. import config

    n, B = prog.n, inp.B
    src = None
    if prog.init[0] in ("input", "amplitude") and inp.state is not None:
        s = inp.state.to(torch.complex64).resolve_conj()
        if s.dim() == 1:
            s = s.unsqueeze(0)
        if s.shape[0] != B:
            s = s.expand(B, -1)
        src = s.contiguous()
    if src is None and efac is not None and checkpoints is None and getattr(config, "ABSORB_LEADING_ROTATIONS", True):
        # the passes then run the remaining gates from the updated factors (same slots, same cs)
        lead = leading_rotations(prog)
        if lead is not None:
            efac = absorb_factors(prog, lead[1], efac, cs, B)
            prog = lead[0]
    jit_on = use_jit(n, R, TB) if specialise is None else specialise
    fold = folds(prog, jit_on) and (not want_state if fold is None else bool(fold))
    pl = get_plan(prog, R, TB, cm, fold)
    ws = pass_words(prog, pl, B, inp.device)
    words = ws.passes
    lib = L.load()
    st = L.stream_of(cs)
    utab, _ = compose_tables(ws, cs, generators=False)
    if out is not None:
        if (out.dtype != torch.complex64 or tuple(out.shape) != (B, 2 ** n) or not out.is_contiguous()
                or out.device != inp.device):
            raise ValueError(f"out must be a contiguous ({B}, {2 ** n}) complex64 tensor on {inp.device}")
        psi = out
    else:
        psi = torch.empty(B, 2 ** n, dtype=torch.complex64, device=inp.device)
    mpart = probs = None
    if prog.measure:
        wb = (lib.qs_readout_work_bytes if fold else lib.qs_fused_measure_work_bytes)(B, n, R, TB)
        mpart = torch.empty(max(wb // 4, 1), dtype=torch.float32, device=inp.device)
    npass = len(words)
    kern = specialised_kernels(prog, pl, ws, B, src is None, fold) if jit_on else None
    ltab = None
    if kern is not None and kern.lanes:
        # operands of the matrix-core unitaries: every table entry (lanes True), or (lanes "lds") each
        # pass's block of the entries it uses, packed pass after pass; none when no pass uses one
        idx = kern.lane_index_on(inp.device) if kern.lanes == "lds" else None
        if kern.lanes is True or idx is not None:
            count = ws.nu if idx is None else idx.numel()
            ltab = torch.empty(max(count, 1) * 16, dtype=torch.float32, device=inp.device)
            L.check(lib.qs_lane_coeffs(L.ptr(ltab), L.ptr(utab), None if idx is None else L.ptr(idx), count, 0,
                                       st), "qs_lane_coeffs")
    ntiles = tiles_of(B, n, R + TB)
    for i, (w, nsub, lpos) in enumerate(words):
        last = i == npass - 1
        store = 1 if (not last or want_state or not prog.measure) else 0
        if events is not None:
            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            ev[0].record(torch.cuda.current_stream(inp.device))
        if kern is not None:
            kern.launch(i, L.ptr(psi), None if src is None else L.ptr(src), L.ptr(utab), L.ptr(cs),
                        None if efac is None else L.ptr(efac),
                        L.ptr(mpart) if (last and mpart is not None) else None, B, ntiles, store, st,
                        None if ltab is None else L.ptr(ltab))
        else:
            L.check(
                lib.qs_fused_pass(
                    L.ptr(psi), None if src is None else L.ptr(src), B, n, R, TB,
                    lpos.ctypes.data_as(L.ctypes.POINTER(L.ctypes.c_int32)), L.ptr(w), nsub,
                    L.ptr(cs), L.ptr(utab), None if efac is None else L.ptr(efac),
                    L.ptr(mpart) if (last and mpart is not None) else None, store, st,
                ),
                "qs_fused_pass",
            )
        if events is not None:
            ev[1].record(torch.cuda.current_stream(inp.device))
            events.append(ev)
        src = psi
        if checkpoints is not None:
            nxt = checkpoints(i, psi)
            if nxt is not None:
                # the caller keeps this pass's output; the next pass reads it and writes nxt
                psi = nxt
    if prog.measure and fold:
        probs = torch.empty(B, n, dtype=torch.float32, device=inp.device)
        L.check(lib.qs_readout_finish(L.ptr(probs), L.ptr(mpart), B, n, R, TB, st), "qs_readout_finish")
    elif prog.measure:
        probs = torch.empty(B, n, dtype=torch.float32, device=inp.device)
        lpos = words[-1][2]
        L.check(
            lib.qs_fused_measure_finish(L.ptr(probs), L.ptr(mpart), B, n, R, TB,
                                        lpos.ctypes.data_as(L.ctypes.POINTER(L.ctypes.c_int32)), st),
            "qs_fused_measure_finish",
        )
    return psi, probs


def run(prog: Program, circuit, state, kwargs):
    """Forward of a lowered circuit (with the adjoint backward when gradients are needed)."""
    from .. import measurements

    inp = gather_inputs(prog, circuit, state, kwargs)
    from . import adjoint

    out_state, probs = adjoint.forward(prog, inp)
    if prog.measure:
        P = probs
        return measurements.reference_shape(P) if not inp.batched else _batched_probs(P)
    out = out_state if inp.batched else out_state.squeeze(0)
    if prog.post == "abs2":
        from .plain import Abs2Fn

        out = Abs2Fn.apply(out)
    return out


def _batched_probs(P: torch.Tensor) -> torch.Tensor:
    from .. import measurements

    return measurements.reference_shape(P)
