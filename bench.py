"""Benchmark of the qandle gate-application hot path on one MI355X.

Headline (the JSON line's ``value``): amplitude-gate updates per second of the
fused forward on the plain-forward workload: 20 qubits, batch 64, complex64,
RX AngleEmbedding (inputs U[0, pi)), 10 layers of RX, RY, RZ on every qubit
(angles U[0, 2 pi)) + CNOT ladder q -> q+1, MeasureProbability.  One step =
one forward of that circuit over the batch, inputs already resident in HBM.
Updates per step = sum over gate applications of B * 2^n (each embedded qubit
counts as one gate).

``roofline`` is the HBM roofline of the dominant kernel, the fused pass kernels
``qs_pass_<i>``: algorithmic bytes per launch (first pass 8 B per amplitude,
middle passes 16 B, the measuring last pass 8 B, over the executed schedule)
divided by the launches' average duration from HIP events on their own
stream, against 8 TB/s.  ``roofline.mfma`` is the same launches against the
fp32 matrix-core peak (16 flops per amplitude per fused 2x2 unitary).

``extras``: the plain per-gate GPU path, 24 qubits x 4, the adjoint forward +
backward at 16 qubits x 256 x 20 layers (with the roofline of its ``qs_adj_<i>``
kernels) against plain per-gate autograd, the 2048-circuit uneven batch (with
the roofline of its level launches), the serving loop, and the 8-qubit CPU
configuration (the reference's dense cached-matrix path restated on the host,
beside the same circuit on the GPU).  ``cpu_baseline``: the structured
complex64 restatement of qandle's gate semantics on the host's CPU share.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick] [--no-cpu]
    python bench.py --profile fused|adjoint|uneven|uneven_train|copy   (short runs for rocprofv3)
    python bench.py --summarize RESULTS_DB [--last N]    (steady-state kernel table)
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_F32_PEAK_TFLOPS = 157.3  # dense fp32 matrix-core peak of one MI355X
_T0 = time.perf_counter()


def log(msg: str) -> None:
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--quick", action="store_true", help="headline only (no extra workloads, no CPU baseline)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", choices=["fused", "adjoint", "uneven", "uneven_train", "copy"],
                    help="run only the named workload a few times (for rocprofv3 counter and trace runs)")
    ap.add_argument("--summarize", metavar="DB", help="print the kernel table of a rocprofv3 results database")
    ap.add_argument("--last", type=int, default=0,
                    help="with --summarize: only the last N dispatches of every kernel (the steady state)")
    return ap.parse_args()


# ----------------------------------------------------------------------------
# workloads
# ----------------------------------------------------------------------------

def headline_workload():
    from qandle_amd import workloads

    return workloads.rxryrz_ladder(n=20, B=64, depth=10, seed=1)


def adjoint_workload():
    from qandle_amd import workloads

    return workloads.rxryrz_ladder(n=16, B=256, depth=20, seed=2, loss=True)


def uneven_workloads():
    import torch

    from qandle_amd import workloads

    ws = workloads.random_uneven(2048, build_layers=False)
    gen = torch.Generator().manual_seed(3)
    return ws, [torch.randn(w.B, w.n, generator=gen) for w in ws]


def settle(fn, sync, seconds=0.3):
    """Run fn (its first call compiles kernels), then repeat it for ``seconds``: after a
    compile the GPU idles for seconds, and the first timing that followed measured up to
    1.5x slower than the same steps a moment later."""
    fn()
    sync()
    t = time.perf_counter()
    while time.perf_counter() - t < seconds:
        fn()
        sync()


def profile_workload(mode, reps=4):
    """Fixed runs for profilers: ``reps`` steady-state steps after the first call and a
    settling period (summarize with --last reps x kernels-per-step)."""
    import torch

    import qandle_amd as qandle

    dev = torch.device("cuda:0")

    def sync():
        torch.cuda.synchronize(dev)

    if mode == "copy":
        n, B = 20, 64
        a = torch.empty(B * 2 ** n, dtype=torch.complex64, device=dev)
        b = torch.empty_like(a)
        for _ in range(reps + 1):
            b.copy_(a)
        sync()
        print(f"copy: {a.numel() * 8} bytes read and written per launch", flush=True)
        return
    if mode in ("uneven", "uneven_train"):
        from qandle_amd._engine import fused
        from qandle_amd._engine.uneven import UnevenBatch

        ws, lw = uneven_workloads()
        ub = UnevenBatch.from_workloads(ws, dev, R=4, TB=8, cost=fused.UNEVEN_COST)
        ub.set_loss_weights(lw)

        def run():
            ub.run_forward()
            if mode == "uneven_train":
                ub.run_backward()

        settle(run, sync)
        for _ in range(reps):
            run()
        sync()
        what = "forwards" if mode == "uneven" else "forward + backward steps"
        print(f"uneven: {reps} {what} of 2048 circuits", flush=True)
        return
    w = headline_workload() if mode == "fused" else adjoint_workload()
    circ = qandle.Circuit(w.layers, num_qubits=w.n).to(dev)
    x = w.inputs["x"].to(dev)
    xg = x.clone().requires_grad_(True)
    lw = None if w.loss_w is None else w.loss_w.to(dev)
    params = list(circ.parameters())

    def step():
        if mode == "fused":
            with torch.no_grad():
                circ(x=x)
        else:  # the adjoint leg's step: gradients cleared as an optimizer does, then forward + backward
            for p in params:
                p.grad = None
            xg.grad = None
            (circ(x=xg) * lw).sum().backward()

    settle(step, sync)
    for _ in range(reps):
        step()
    sync()
    print(f"{mode}: {reps} steady-state steps of {w.name}", flush=True)


def summarize(db: str, last: int):
    """Kernel table (calls, total, average, min, max in microseconds) of a rocprofv3
    results database; with ``last`` > 0 only the last ``last`` dispatches of every
    kernel name count (the steady-state steps after compiles and settling).  A
    ``*counter_collection.csv`` file instead gives, per kernel and counter, the
    average value per dispatch over the same last dispatches."""
    if db.endswith(".csv"):
        summarize_counters(db, last)
        return
    import sqlite3

    con = sqlite3.connect(db)
    rows = con.execute("select name, start, end from kernels order by start").fetchall()
    by = {}
    for name, s, e in rows:
        by.setdefault(name, []).append((e - s) / 1e3)
    print(f"| Kernel | Calls | Total (us) | Average (us) | Min (us) | Max (us) |")
    print("|---|---|---|---|---|---|")
    out = []
    for name, ds in by.items():
        if last > 0:
            ds = ds[-last:]
        out.append((sum(ds), name, ds))
    for tot, name, ds in sorted(out, reverse=True):
        short = name if len(name) < 70 else name[:67] + "..."
        print(f"| `{short}` | {len(ds)} | {tot:.1f} | {tot / len(ds):.1f} | {min(ds):.1f} | {max(ds):.1f} |")


def summarize_counters(path: str, last: int):
    import csv

    per = {}  # (kernel, counter) -> {dispatch: value}
    with open(path) as f:
        for row in csv.DictReader(f):
            key = (row["Kernel_Name"], row["Counter_Name"])
            d = per.setdefault(key, {})
            did = int(row["Dispatch_Id"])
            d[did] = d.get(did, 0.0) + float(row["Counter_Value"])
    print("| Kernel | Counter | Dispatches | Average per dispatch |")
    print("|---|---|---|---|")
    for (kern, ctr), d in sorted(per.items()):
        vals = [d[k] for k in sorted(d)]
        if last > 0:
            vals = vals[-last:]
        short = kern if len(kern) < 70 else kern[:67] + "..."
        print(f"| `{short}` | {ctr} | {len(vals)} | {sum(vals) / len(vals):.6g} |")


# ----------------------------------------------------------------------------
# helpers for the headline roofline
# ----------------------------------------------------------------------------

def executed_program(prog):
    """The program a forward without gradients runs from a product state: the
    rotations ahead of every CNOT on their qubit go into the qubits' factors
    (fused.leading_rotations)."""
    from qandle_amd import config
    from qandle_amd._engine import fused

    if config.ABSORB_LEADING_ROTATIONS and fused.leading_rotations(prog) is not None:
        return fused.leading_rotations(prog)[0]
    return prog


def executed_plan(prog, R, TB):
    """The plan of executed_program, with the compiled kernels' readout fold of
    the trailing CNOTs (fused.readout_fold)."""
    from qandle_amd._engine import fused

    p = executed_program(prog)
    return fused.get_plan(p, R, TB, None, fused.folds(p, fused.use_jit(p.n, R, TB)))


def pass_structure(prog, plan, B):
    """(fused 2x2 unitaries, sub-passes) of every pass of the executed schedule."""
    from qandle_amd._engine import fused
    from qandle_amd._engine import planner as PL

    ws = fused.pass_words(executed_program(prog), plan, B, "cpu")
    out = []
    for w, nsub, _lp in ws.passes:
        subs = PL.decode_pass(w.numpy(), plan.R, plan.TB)
        out.append((sum(len(bits) for sub in subs for bits, _cn in sub.segments), nsub))
    return out


def pass_bytes(B, n, npass, stores_last):
    """Algorithmic HBM bytes of every pass launch: the first pass generates the
    embedding product state in registers and writes it (8 B per amplitude), middle
    passes read and write (16 B), the last pass reads (8 B, plus a write when the
    state is kept).  Measurement partials are K + 1 (or n + 1) floats per tile."""
    amps = B * 2 ** n
    if npass == 1:
        return [amps * (8 if stores_last else 0)]
    return [amps * 8] + [amps * 16] * (npass - 2) + [amps * (16 if stores_last else 8)]


MFMA_PROBE = r"""
typedef float qs_f4 __attribute__((ext_vector_type(4)));
extern "C" __global__ void __launch_bounds__(256) qs_mfma_probe(float *out, int iters, float a0) {
  qs_f4 c[8];
  for (int k = 0; k < 8; ++k) c[k] = qs_f4{0.f, 0.f, 0.f, 0.f};
  const float a = a0 + threadIdx.x * 1e-7f, b = a0 * 0.5f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c[k], 0, 0, 0);
  }
  float s = 0.f;
  for (int k = 0; k < 8; ++k) s += c[k].x + c[k].y + c[k].z + c[k].w;
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
"""


def mfma_sustained(dev, stream):
    """The fp32 matrix-core rate this GPU sustains, measured: 8 independent chains
    of v_mfma_f32_4x4x1 (the pass kernels' instruction) per wave, with 2 and with
    8 waves per SIMD (2 is the pass kernels' occupancy).  The published 157.3
    TFLOP/s assumes the peak engine clock; under matrix-core load the clock runs
    lower, and this is the rate the pass kernels can approach."""
    import ctypes
    import struct as st

    import torch

    from qandle_amd._engine import jit, lib as L

    fn = jit.kernels_for([MFMA_PROBE], ["qs_mfma_probe"])[0]
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    out = {}
    for waves in (2, 8):
        blocks = waves * cus                      # 4 waves per 256-thread block, 4 SIMDs per CU
        iters = max(64, int(15000 * 2 / waves))
        buf_t = torch.empty(blocks * 256, dtype=torch.float32, device=dev)
        best = None
        for rep in range(4):
            args = st.pack("<Qif", buf_t.data_ptr(), iters if rep else 16, 1.0)
            buf = ctypes.create_string_buffer(args, len(args))
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            L.check(L.load().qs_jit_launch(fn, blocks, 256, buf, len(args), stream.cuda_stream), "qs_jit_launch")
            e1.record(stream)
            torch.cuda.synchronize(dev)
            if rep:
                t = e0.elapsed_time(e1) / 1e3
                best = t if best is None else min(best, t)
        out[f"{waves}_waves_per_simd"] = blocks * 4 * iters * 8 * 512 / best / 1e12
    return out


def hbm_roofline(events, kernel):
    """HBM roofline dict from (start, end, bytes) launch records."""
    durs = [a.elapsed_time(b) / 1e3 for a, b, _ in events]
    nbytes = sum(x for _, _, x in events)
    achieved = nbytes / sum(durs) / 1e9
    return {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "avg_launch_us": sum(durs) / len(durs) * 1e6,
            "alg_bytes_per_launch": nbytes / len(durs), "launches_timed": len(durs)}


def cpu_share() -> int:
    """Host CPUs this process may use: its affinity set, capped by OMP_NUM_THREADS
    when the environment sets it (a host's CPU share per GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(n, 1)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def median_time(fn, runs=5):
    fn()  # warm-up
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts), ts


# ----------------------------------------------------------------------------
# main
# ----------------------------------------------------------------------------

def main():
    args = parse()
    if args.summarize:
        summarize(args.summarize, args.last)
        return
    if args.profile:
        profile_workload(args.profile)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine runs on a single MI355X; "
              "sharding a state across GPUs is not implemented.", file=sys.stderr)
        sys.exit(2)

    warm = (0, 0.0)
    if not args.quick:
        log("serving mix: compiling its program structures in worker processes (before any GPU use)")
        warm = serving_warmup()
        log(f"serving mix: {warm[0]} program structures compiled in {warm[1]:.1f} s")

    import torch

    import qandle_amd as qandle
    from qandle_amd._engine import fused, lib

    lib.load()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)

    def sync():
        torch.cuda.synchronize(dev)

    def time_steps(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        sync()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        sync()
        wall = time.perf_counter() - t0
        return e0.elapsed_time(e1) / 1e3, wall

    # ---------------------------------------------------------------- headline
    log("headline: 20 qubits, batch 64, 10 layers (first call compiles the pass kernels)")
    w = headline_workload()
    n, B, depth = w.n, w.B, 10
    circ = qandle.Circuit(w.layers, num_qubits=n).to(dev)
    x = w.inputs["x"].to(dev)
    prog = fused.lower_circuit(circ.circuit)
    R, TB = fused.choose_shape(n)

    def step():
        with torch.no_grad():
            return circ(x=x)

    settle(step, sync)
    gpu_s, wall_s = time_steps(step, args.steps, args.warmup)
    ms_per_step = gpu_s / args.steps * 1e3
    log(f"headline: {ms_per_step:.3f} ms per step")
    updates = w.amplitude_updates()
    value = updates * args.steps / gpu_s

    # per-launch durations of the dominant kernels, on their own stream
    plan = executed_plan(prog, R, TB)
    npass = len(plan.passes)
    pbytes = pass_bytes(B, n, npass, stores_last=False)
    struct = pass_structure(prog, plan, B)
    with torch.no_grad():
        inp = fused.gather_inputs(prog, circ.circuit, None, {"x": x})
        cs = fused.build_coeffs(prog, inp)
        efac = fused.product_factors(prog, inp)
        raw = []
        reps = max(2, args.steps // 4)
        for _ in range(reps):
            fused.execute(prog, inp, cs, efac, R, TB, want_state=False, events=raw)
        sync()
    events = [(a, b, pbytes[i % npass]) for i, (a, b) in enumerate(raw)]
    kernel = (f"qs_pass_<i> (pass kernels specialised per pass program, hipRTC; R={R}, TB={TB})"
              if fused.use_jit(n, R, TB) else f"k_pass<{R},{TB},false> (qs_fused_pass)")
    roof = hbm_roofline(events, kernel)
    per_pass = []
    for i in range(npass):
        ds = [a.elapsed_time(b) * 1e3 for a, b, _ in events[i::npass]]
        per_pass.append({"us": sum(ds) / len(ds), "unitaries": struct[i][0], "sub_passes": struct[i][1],
                         "alg_bytes": pbytes[i]})
    units = sum(u for u, _ in struct)
    flops_per_launch = units / npass * B * 2 ** n * 16
    tflops = flops_per_launch / (roof["avg_launch_us"] / 1e6) / 1e12
    sustained = mfma_sustained(dev, stream)
    roof["mfma"] = {"achieved": tflops, "peak": MFMA_F32_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": tflops / MFMA_F32_PEAK_TFLOPS, "alg_flops_per_launch": flops_per_launch,
                    "fused_unitaries_per_step": units,
                    "sustained_measured": sustained,
                    "frac_of_sustained": tflops / sustained["2_waves_per_simd"]}
    roof["per_pass"] = per_pass
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_fused_pass_n20_b64.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roof["traffic"] = traffic

    result = {
        "metric": "amplitude-gate updates/s, % HBM peak, fused vs plain; p50/p99 serve latency",
        "value": value,
        "unit": "amplitude-gate updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded inputs and angles, remapping=None)",
        "config": {
            "workload": "plain forward: 20 qubits, batch 64, RX AngleEmbedding + 10 x (RX,RY,RZ per qubit + CNOT ladder) + MeasureProbability",
            "qubits": n, "batch": B, "layers": depth, "gate_ops": w.gate_ops,
            "updates_per_step": updates, "fused_passes": npass, "tile": {"R": R, "TB": TB},
            "parallelism": "single GPU",
        },
        "roofline": roof,
        "wall_s_timed_region": wall_s,
    }

    if not args.quick:
        extras = {}
        log("plain per-gate path on the headline workload")

        def plain_step():
            with torch.no_grad():
                return circ.circuit.forward_layers(x=x)
        p_s, _ = time_steps(plain_step, 2, 1)
        plain_bytes = B * 2 ** n * (8 + 16 * len(w.gates) + 8)
        extras["plain_updates_per_s"] = updates * 2 / p_s
        extras["plain_ms_per_step"] = p_s / 2 * 1e3
        extras["plain_step_hbm_frac"] = plain_bytes / (p_s / 2) / 1e9 / HBM_PEAK_GBS
        extras["fused_over_plain"] = extras["plain_ms_per_step"] / ms_per_step
        extras["fused_alg_bytes_per_step"] = sum(pbytes)
        extras["fused_step_hbm_frac"] = sum(pbytes) / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS

        log("24 qubits, batch 4")
        extras.update(bench_n24(dev, time_steps, sync))
        torch.cuda.empty_cache()
        log("adjoint forward + backward: 16 qubits, batch 256, 20 layers")
        extras.update(bench_adjoint(dev, time_steps, sync))
        torch.cuda.empty_cache()
        log("uneven batch of 2048 circuits")
        extras.update(bench_uneven(dev, time_steps))
        torch.cuda.empty_cache()
        log("serving loop")
        extras.update(bench_serving(dev, warm))
        torch.cuda.empty_cache()
        if not args.no_cpu:
            log("8-qubit configuration: dense reference path on the CPU beside the GPU forward")
            extras.update(bench_c1(dev, time_steps, sync))
        result["extras"] = extras
        if not args.no_cpu:
            log("CPU baseline sample")
            result["cpu_baseline"] = cpu_baseline(w)
    log("done")
    print(json.dumps(result))


def bench_n24(dev, time_steps, sync):
    import torch

    import qandle_amd as qandle
    from qandle_amd import workloads
    from qandle_amd._engine import fused

    w = workloads.rxryrz_ladder(n=24, B=4, depth=10, seed=1)
    c = qandle.Circuit(w.layers, num_qubits=24).to(dev)
    x = w.inputs["x"].to(dev)

    def fwd():
        with torch.no_grad():
            c(x=x)
    settle(fwd, sync)
    t, _ = time_steps(fwd, 5, 2)
    prog = fused.lower_circuit(c.circuit)
    npass = len(executed_plan(prog, *fused.choose_shape(24)).passes)
    nbytes = sum(pass_bytes(4, 24, npass, stores_last=False))
    return {"n24_updates_per_s": w.amplitude_updates() * 5 / t, "n24_ms_per_step": t / 5 * 1e3,
            "n24_hbm_frac": nbytes / (t / 5) / 1e9 / HBM_PEAK_GBS, "n24_passes": npass}


def bench_adjoint(dev, time_steps, sync):
    import torch

    import qandle_amd as qandle
    from qandle_amd._engine import backward

    w = adjoint_workload()
    c3 = qandle.Circuit(w.layers, num_qubits=w.n).to(dev)
    x3 = w.inputs["x"].to(dev).requires_grad_(True)
    lw = w.loss_w.to(dev)
    # gradients are cleared as an optimizer's zero_grad(set_to_none=True) does, over the
    # parameter list taken once (Module.zero_grad walks the 960 gate modules every call)
    params = list(c3.parameters())

    def clear():
        for p in params:
            p.grad = None
        x3.grad = None

    def adj_step():
        clear()
        (c3(x=x3) * lw).sum().backward()

    def plain_adj_step():
        clear()
        (c3.circuit.forward_layers(x=x3) * lw).sum().backward()

    settle(adj_step, sync)
    blocks = sorted(time_steps(adj_step, 10, 2)[0] / 10 for _ in range(3))
    ta = blocks[1] * 6  # the median block, per 6 steps (below)
    log(f"adjoint: {blocks[1] * 1e3:.2f} ms per forward + backward (median of 3 blocks of 10 steps); "
        "adjoint kernel launches, then plain autograd")
    events = []
    backward.launch_events = events
    try:
        for _ in range(2):
            adj_step()
        sync()
    finally:
        backward.launch_events = None
    roof = hbm_roofline(events, "qs_adj_<i> (adjoint sweep steps specialised per step program, hipRTC)")
    tp, _ = time_steps(plain_adj_step, 1, 1)
    out = {"adjoint_fwd_bwd_ms": ta / 6 * 1e3, "adjoint_fwd_bwd_blocks_ms": [b * 1e3 for b in blocks],
           "plain_autograd_fwd_bwd_ms": tp * 1e3,
           "adjoint_speedup_vs_plain": tp / (ta / 6), "adjoint_roofline": roof}
    del c3, x3
    return out


def bench_uneven(dev, time_steps):
    import torch

    from qandle_amd._engine import autotune
    from qandle_amd._engine.uneven import UnevenBatch

    ws, lw = uneven_workloads()

    def make(R, TB, cost):
        ub = UnevenBatch.from_workloads(ws, dev, R=R, TB=TB, cost=cost)
        ub.set_loss_weights(lw)
        return ub

    ub, table = autotune.tune_uneven(make, cands=((4, 8, 1024.0), (4, 8, 2048.0), (4, 7, 1024.0)), reps=2)
    ub.capture(backward=False)
    tf, _ = time_steps(lambda: ub.replay(backward=False), 5, 1)
    tfb, _ = time_steps(lambda: ub.replay(backward=True), 5, 1)
    te, _ = time_steps(lambda: ub.run_forward(), 3, 1)
    events = []
    for _ in range(2):
        ub.run_forward(events=events)
    torch.cuda.synchronize(dev)
    roof = hbm_roofline(events, f"qs_items_pass (level launches of (circuit pass, tile) work items; R={ub.R}, TB={ub.TB})")
    upd = ub.amplitude_updates
    return {
        "uneven_circuits": len(ws),
        "uneven_updates_per_fwd": upd,
        "uneven_fwd_ms": tf / 5 * 1e3,
        "uneven_fwd_updates_per_s": upd * 5 / tf,
        "uneven_fwd_bwd_ms": tfb / 5 * 1e3,
        "uneven_fwd_eager_ms": te / 3 * 1e3,
        "uneven_levels": len(ub.fw_level_items),
        "uneven_autotune": table,
        "uneven_roofline": roof,
    }


SERVE_COMPILED_FROM = 18  # requests of this many qubits and more run compiled kernels (99.8% of the mix's work)
SERVE_FIXED_RATE_HZ = 7.5  # a fixed offered load of the request mix, for comparing serving runs


SERVE_MIX = dict(num=160, seed=4)


def serving_warmup():
    """Compile the serving mix's program structures in worker processes, into the
    code-object cache.  Runs before this process touches the GPU: the workers are
    started as fresh interpreters and only compile (hipRTC), they launch nothing."""
    from qandle_amd._engine import serving

    t0 = time.perf_counter()
    reqs = serving.poisson_requests(SERVE_MIX["num"], 1.0, SERVE_MIX["seed"], "cpu")
    structures = serving.precompile(reqs, SERVE_COMPILED_FROM, workers=cpu_share())
    return structures, time.perf_counter() - t0


def bench_serving(dev, warm, num=SERVE_MIX["num"], seed=SERVE_MIX["seed"]):
    """Open-loop serving at 70% of the serving path's own capacity on this
    request mix (measured first by a closed-loop run of the same requests, all
    arriving at once); reports p50/p99 latency from arrival to completion and
    p50/p99 service time from admission to completion.  The compiled kernels of
    the mix's program structures are built before the serving clock starts, as a
    server warms its kernel cache for its circuit templates at start-up."""
    from qandle_amd._engine import serving

    def fresh():
        reqs = serving.poisson_requests(num, 1.0, seed, dev)
        eng = serving.ServingEngine(dev, cap_bytes=200e9, streams=4, compiled_from_qubits=SERVE_COMPILED_FROM)
        for r in reqs:
            eng.prepare(r)
        return reqs, eng

    reqs, eng = fresh()
    work = sum((len(serving.fused.lower_circuit(r.circuit.circuit).gates) + r.n) * r.B * 2 ** r.n for r in reqs)
    arrivals = [r.arrival for r in reqs]
    for r in reqs:
        r.arrival = 0.0
    t0 = time.perf_counter()
    eng.serve(reqs, timeout_s=300)
    busy = time.perf_counter() - t0             # closed-loop time of the whole mix
    log(f"serving: closed-loop {busy:.1f} s for {num} requests")
    reqs, eng = fresh()
    rate = 0.7 * num / busy                     # 70% offered load
    for r, a in zip(reqs, arrivals):
        r.arrival = a / rate                    # arrivals were drawn at rate 1
    t0 = time.perf_counter()
    done = eng.serve(reqs, timeout_s=300)
    wall = time.perf_counter() - t0
    st = serving.latency_stats(done)
    # the same requests at a fixed offered load, so that runs compare at the same load
    reqs, eng = fresh()
    for r, a in zip(reqs, arrivals):
        r.arrival = a / SERVE_FIXED_RATE_HZ
    fixed = serving.latency_stats(eng.serve(reqs, timeout_s=300))
    return {
        "serve_requests": st["requests"],
        "serve_rejected": st["rejected"],
        "serve_p50_ms": st["p50_ms"],
        "serve_p99_ms": st["p99_ms"],
        "serve_service_p50_ms": st["service_p50_ms"],
        "serve_service_p99_ms": st["service_p99_ms"],
        "serve_compiled_from_qubits": SERVE_COMPILED_FROM,
        "serve_structures_compiled": warm[0],
        "serve_kernel_warmup_s": warm[1],
        "serve_rate_hz": rate,
        "serve_closed_loop_s": busy,
        "serve_wall_s": wall,
        "serve_updates_per_s": work / wall,
        "serve_peak_resident_gb": eng.peak_resident / 1e9,
        "serve_max_qubits": int(max(r.n for r in reqs)),
        "serve_fixed_rate_hz": SERVE_FIXED_RATE_HZ,
        "serve_fixed_p50_ms": fixed["p50_ms"],
        "serve_fixed_p99_ms": fixed["p99_ms"],
        "serve_fixed_service_p50_ms": fixed["service_p50_ms"],
    }


def bench_c1(dev, time_steps, sync):
    """The 8-qubit configuration (batch 32, RX embedding, 4 x (RY per qubit + CNOT
    ring), MeasureProbability): the reference's own cached-matrix path restated in
    complex64 torch on the host (dense 2^n x 2^n gate matrices, the embedding's
    einsum + matmul chain; oracle/qandle_c64.py, dense mode), median of 5 forwards
    on the host's CPU share after the cache build, beside the fused GPU forward."""
    import torch

    import qandle_amd as qandle
    from oracle import qandle_c64
    from qandle_amd import workloads

    w = workloads.embed_ry_ring(n=8, B=32, depth=4, seed=0)
    threads = cpu_share()
    torch.set_num_threads(threads)
    ops = w.torch_ops()
    t0 = time.perf_counter()
    dense = qandle_c64.DenseCircuit(w.n, ops)
    build_s = time.perf_counter() - t0
    med, runs = median_time(lambda: dense(ops))
    c = qandle.Circuit(w.layers, num_qubits=w.n).to(dev)
    x = w.inputs["x"].to(dev)

    def fwd():
        with torch.no_grad():
            c(x=x)
    settle(fwd, sync)
    tg, _ = time_steps(fwd, 50, 5)
    upd = w.amplitude_updates()
    return {"c1_workload": "8 qubits, batch 32, RX embedding + 4 x (RY per qubit + CNOT ring) + MeasureProbability",
            "c1_cpu_dense_ms": med * 1e3, "c1_cpu_dense_runs_ms": [t * 1e3 for t in runs],
            "c1_cpu_dense_build_s": build_s, "c1_cpu_threads": threads,
            "c1_cpu_dense_updates_per_s": upd / med,
            "c1_gpu_ms": tg / 50 * 1e3, "c1_gpu_updates_per_s": upd * 50 / tg,
            "c1_gpu_over_cpu_dense": med / (tg / 50)}


def cpu_baseline(w):
    """qandle's complex64 torch semantics restated (oracle/qandle_c64.py, structured
    mode: the dense reference needs 2 x 8 TiB of cached matrices per gate at 20
    qubits) on a bounded sample of the headline workload: the embedding and the
    first 3 of the 10 layers (257 of the 810 gate applications) on 12 of the 64
    samples, median of 3 runs after a warm-up (20-30 s of CPU work), on every CPU of
    this process's share.  The sample's state (100 MB) is past the host's caches, as
    the full batch's (537 MB) is: a 4-sample state (34 MB) stays cache-resident and
    runs several times faster per amplitude than the workload would."""
    import torch

    from oracle import qandle_c64

    threads = cpu_share()
    torch.set_num_threads(threads)
    Bs, layers = 12, 3
    per_layer = (len(w.gates)) // 10
    ops = w.torch_ops()[: 1 + layers * per_layer]
    ops[0] = (ops[0][0], ops[0][1], ops[0][2], ops[0][3][:Bs])
    gates = w.n + layers * per_layer  # each embedded qubit counts as one gate, as in the headline
    med, runs = median_time(lambda: qandle_c64.structured_run(w.n, ops, Bs), runs=3)
    updates = gates * Bs * 2 ** w.n
    return {
        "value": updates / med,
        "unit": "amplitude-gate updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"embedding + first {layers} of 10 layers ({gates} of {w.gate_ops} gate applications) on {Bs} "
                  f"of the {w.B} samples, torch CPU complex64 (structured restatement of qandle's gate "
                  f"semantics), median of 3 runs ({', '.join(f'{t:.2f}' for t in runs)} s)",
        "cpu": _cpu_model(),
        "host_cpus": os.cpu_count(),
        "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
    }


if __name__ == "__main__":
    main()
