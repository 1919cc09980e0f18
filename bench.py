"""Benchmark of the qandle gate-application hot path on one MI355X.

Headline (the JSON line's ``value``): amplitude-gate updates per second of the
fused forward on the plain-forward workload: 20 qubits, batch 64, complex64,
RX AngleEmbedding (inputs U[0, pi)), 10 layers of RX, RY, RZ on every qubit
(angles U[0, 2 pi)) + CNOT ladder q -> q+1, MeasureProbability.  One step =
one forward of that circuit over the batch, inputs already resident in HBM.
Updates per step = sum over gate applications of B * 2^n (each embedded qubit
counts as one gate).

Also reported: the plain per-gate GPU path on the same workload, the fused
forward at 24 qubits (batch 4), the adjoint forward + backward at 16 qubits
(batch 256, 20 layers) against the plain per-gate path, the roofline of the
dominant kernel (the fused pass kernel), and the CPU restatement of qandle's
complex64 torch path timed on a bounded sample on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_F32_PEAK_TFLOPS = 157.3  # dense fp32 matrix-core peak of one MI355X
CPU_THREADS = 16
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
    ap.add_argument("--profile", choices=["fused", "adjoint", "copy"],
                    help="run only the named workload a few times (for rocprofv3 counter and trace runs)")
    return ap.parse_args()


def profile_workload(mode, reps=2):
    """Short fixed runs for profilers: the headline forward, its forward + adjoint
    backward, or a device copy of the same state size (the HBM counter calibration)."""
    import torch

    import qandle_amd as qandle
    from qandle_amd import workloads

    dev = torch.device("cuda:0")
    n, B = 20, 64
    if mode == "copy":
        a = torch.empty(B * 2 ** n, dtype=torch.complex64, device=dev)
        b = torch.empty_like(a)
        for _ in range(reps + 1):
            b.copy_(a)
        torch.cuda.synchronize(dev)
        print(f"copy: {a.numel() * 8} bytes read and written per launch", flush=True)
        return
    w = workloads.rxryrz_ladder(n=n, B=B, depth=10, seed=1)
    circ = qandle.Circuit(w.layers, num_qubits=n).to(dev)
    x = w.inputs["x"].to(dev)
    for _ in range(reps):
        if mode == "fused":
            with torch.no_grad():
                circ(x=x)
        else:
            xg = x.clone().requires_grad_(True)
            circ(x=xg).sum().backward()
    torch.cuda.synchronize(dev)
    print(f"{mode}: {reps} runs of {w.name}", flush=True)


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


def fused_unitaries(prog, plan, B):
    """Fused 2x2 unitaries the executed schedule applies to every amplitude, all passes together."""
    from qandle_amd._engine import fused
    from qandle_amd._engine import planner as PL

    ws = fused.pass_words(executed_program(prog), plan, B, "cpu")
    return sum(len(bits) for w, _ns, _lp in ws.passes for sub in PL.decode_pass(w.numpy(), plan.R, plan.TB)
               for bits, _cn in sub.segments)


def fused_bytes_per_launch(B, n, npass, stores_last):
    """Algorithmic HBM bytes of the executed schedule, averaged per pass launch.

    First pass: generates the embedding product state in registers and writes
    it (8 B/amplitude).  Middle passes: read + write (16 B).  Last pass: read
    (8 B), plus a write when the state is kept.  Measurement partials are
    K + 1 (or, with a folded readout, n + 1) floats per tile, negligible.
    """
    amps = B * 2 ** n
    per_amp = 8 + 16 * max(npass - 2, 0) + (8 + (8 if stores_last else 0) if npass > 1 else 0)
    if npass == 1:
        per_amp = 8 if stores_last else 0
    return amps * per_amp / npass, amps * per_amp


def main():
    args = parse()
    if args.profile:
        profile_workload(args.profile)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine runs on a single MI355X; "
              "sharding a state across GPUs is not implemented.", file=sys.stderr)
        sys.exit(2)

    import torch

    import qandle_amd as qandle
    from qandle_amd import workloads
    from qandle_amd._engine import fused, lib

    lib.load()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)

    def sync():
        torch.cuda.synchronize(dev)

    def settle(fn, seconds=0.3):
        """Run fn (its first call compiles kernels) and then repeat it for ``seconds``: after a
        compile the GPU idles for seconds, and the first timing that follows measured up to
        1.5x slower than the same steps a moment later (25.7 vs 17.2 ms for the adjoint)."""
        fn()
        sync()
        t = time.perf_counter()
        while time.perf_counter() - t < seconds:
            fn()
            sync()

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
    n, B, depth = 20, 64, 10
    w = workloads.rxryrz_ladder(n=n, B=B, depth=depth, seed=1)
    circ = qandle.Circuit(w.layers, num_qubits=n).to(dev)
    x = w.inputs["x"].to(dev)
    prog = fused.lower_circuit(circ.circuit)
    R, TB = fused.choose_shape(n)

    def step():
        with torch.no_grad():
            return circ(x=x)

    # first call (compiles the pass kernels), then ~0.3 s of the same steps before the W warm-up
    # steps (the first of several timings in a process measured ~10% slower than the rest)
    settle(step)
    gpu_s, wall_s = time_steps(step, args.steps, args.warmup)
    ms_per_step = gpu_s / args.steps * 1e3
    log(f"headline: {ms_per_step:.2f} ms per step")
    updates = w.amplitude_updates()
    value = updates * args.steps / gpu_s

    # per-launch durations of the dominant kernel, on its own stream, over K steps
    plan = executed_plan(prog, R, TB)
    npass = len(plan.passes)
    with torch.no_grad():
        inp = fused.gather_inputs(prog, circ.circuit, None, {"x": x})
        cs = fused.build_coeffs(prog, inp)
        efac = fused.product_factors(prog, inp)
        events = []
        for _ in range(max(2, args.steps // 4)):
            fused.execute(prog, inp, cs, efac, R, TB, want_state=False, events=events)
        sync()
    durs = [a.elapsed_time(b) / 1e3 for a, b in events]
    avg_launch = sum(durs) / len(durs)
    bpl, bstep = fused_bytes_per_launch(B, n, npass, stores_last=False)
    achieved = bpl / avg_launch / 1e9
    # matrix-core work: each fused unitary is a 2x2 complex matvec per amplitude pair,
    # 16 real multiply-adds = 32 flops per pair, 16 per amplitude
    units = fused_unitaries(prog, plan, B)
    fpl = units / npass * B * 2 ** n * 16
    tflops = fpl / avg_launch / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_fused_pass_n20_b64.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

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
        "roofline": {
            # the pass kernels carry ~26 fused unitaries each: the fp32 matrix-core rate is the
            # nearer roof (the HBM figures of the same launches are kept beside it)
            "bound": "mfma",
            "kernel": (f"qs_pass_<i> (pass kernels specialised per pass program, hipRTC; R={R}, TB={TB})"
                       if fused.use_jit(n, R, TB) else f"k_pass<{R},{TB},false> (qs_fused_pass)"),
            "achieved": tflops,
            "peak": MFMA_F32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": tflops / MFMA_F32_PEAK_TFLOPS,
            "traffic": traffic,
            "avg_launch_us": avg_launch * 1e6,
            "alg_flops_per_launch": fpl,
            "fused_unitaries_per_step": units,
            "hbm": {"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                    "alg_bytes_per_launch": bpl},
            "launches_timed": len(durs),
        },
        "wall_s_timed_region": wall_s,
    }

    if not args.quick:
        extras = {}
        # plain per-gate GPU path on the same workload
        log("plain per-gate path on the headline workload")

        def plain_step():
            with torch.no_grad():
                return circ.circuit.forward_layers(x=x)
        p_s, _ = time_steps(plain_step, 2, 1)
        extras["plain_updates_per_s"] = updates * 2 / p_s
        extras["plain_ms_per_step"] = p_s / 2 * 1e3
        extras["fused_over_plain"] = extras["plain_ms_per_step"] / ms_per_step
        extras["fused_alg_bytes_per_step"] = bstep
        extras["fused_step_hbm_frac"] = bstep / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS
        extras["plain_step_hbm_frac"] = (B * 2 ** n * (8 + 16 * len(w.gates) + 8)) / (p_s / 2) / 1e9 / HBM_PEAK_GBS
        # the bytes a gate-by-gate streaming path moves, over the fused step's time
        extras["fused_per_gate_equivalent_hbm_frac"] = ((B * 2 ** n * (8 + 16 * len(w.gates) + 8)) / (ms_per_step / 1e3)
                                                        / 1e9 / HBM_PEAK_GBS)

        # 24 qubits, batch 4 (same circuit family)
        log("24 qubits, batch 4")
        w24 = workloads.rxryrz_ladder(n=24, B=4, depth=10, seed=1)
        c24 = qandle.Circuit(w24.layers, num_qubits=24).to(dev)
        x24 = w24.inputs["x"].to(dev)
        with torch.no_grad():
            settle(lambda: c24(x=x24))
            t24, _ = time_steps(lambda: c24(x=x24), 5, 2)
        p24 = fused.lower_circuit(c24.circuit)
        np24 = len(executed_plan(p24, *fused.choose_shape(24)).passes)
        _, b24 = fused_bytes_per_launch(4, 24, np24, stores_last=False)
        extras["n24_updates_per_s"] = w24.amplitude_updates() * 5 / t24
        extras["n24_ms_per_step"] = t24 / 5 * 1e3
        extras["n24_hbm_frac"] = b24 / (t24 / 5) / 1e9 / HBM_PEAK_GBS
        extras["n24_per_gate_equivalent_hbm_frac"] = (4 * 2 ** 24 * (8 + 16 * len(w24.gates) + 8)) / (t24 / 5) / 1e9 / HBM_PEAK_GBS
        extras["n24_passes"] = np24

        # adjoint forward + backward, 16 qubits, batch 256, 20 layers
        log("adjoint forward + backward: 16 qubits, batch 256, 20 layers")
        w3 = workloads.rxryrz_ladder(n=16, B=256, depth=20, seed=2, loss=True)
        c3 = qandle.Circuit(w3.layers, num_qubits=16).to(dev)
        x3 = w3.inputs["x"].to(dev).requires_grad_(True)
        lw = w3.loss_w.to(dev)

        def adj_step():
            c3.zero_grad(set_to_none=True)
            x3.grad = None
            (c3(x=x3) * lw).sum().backward()

        def plain_adj_step():
            c3.zero_grad(set_to_none=True)
            x3.grad = None
            (c3.circuit.forward_layers(x=x3) * lw).sum().backward()

        settle(adj_step)
        ta, _ = time_steps(adj_step, 6, 2)
        log(f"adjoint: {ta / 6 * 1e3:.2f} ms per forward + backward; plain per-gate autograd next")
        tp, _ = time_steps(plain_adj_step, 1, 1)
        extras["adjoint_fwd_bwd_ms"] = ta / 6 * 1e3
        extras["plain_autograd_fwd_bwd_ms"] = tp * 1e3
        extras["adjoint_speedup_vs_plain"] = tp / (ta / 6)
        del c3, x3
        torch.cuda.empty_cache()

        # uneven batch: 2048 circuits, 4..18 qubits, depth 2..24, forward + adjoint in one graph
        log("uneven batch of 2048 circuits")
        extras.update(bench_uneven(dev, time_steps))
        torch.cuda.empty_cache()
        # serving: Poisson arrivals at ~70% offered load of the measured fused throughput
        log("serving loop")
        extras.update(bench_serving(dev))
        torch.cuda.empty_cache()
        result["extras"] = extras

        if not args.no_cpu:
            log("CPU baseline sample")
            result["cpu_baseline"] = cpu_baseline(w)
    log("done")
    print(json.dumps(result))


def bench_uneven(dev, time_steps):
    import torch

    from qandle_amd import workloads
    from qandle_amd._engine import autotune
    from qandle_amd._engine.uneven import UnevenBatch

    ws = workloads.random_uneven(2048, build_layers=False)
    gen = torch.Generator().manual_seed(3)
    lw = [torch.randn(w.B, w.n, generator=gen) for w in ws]

    def make(R, TB, cost):
        ub = UnevenBatch.from_workloads(ws, dev, R=R, TB=TB, cost=cost)
        ub.set_loss_weights(lw)
        return ub

    ub, table = autotune.tune_uneven(make, cands=((4, 7, 160.0), (4, 8, 160.0)), reps=2)
    ub.capture(backward=False)
    tf, _ = time_steps(lambda: ub.replay(backward=False), 5, 1)
    tfb, _ = time_steps(lambda: ub.replay(backward=True), 5, 1)
    te, _ = time_steps(lambda: ub.run_forward(), 3, 1)
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
    }


def bench_serving(dev, num=160, seed=4):
    """Open-loop serving at 70% of the serving path's own capacity on this
    request mix (measured first by a closed-loop run of the same requests, all
    arriving at once); reports p50/p99 latency from arrival to completion."""
    from qandle_amd._engine import serving

    def fresh():
        reqs = serving.poisson_requests(num, 1.0, seed, dev)
        eng = serving.ServingEngine(dev, cap_bytes=200e9, streams=4)
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
    return {
        "serve_requests": st["requests"],
        "serve_rejected": st["rejected"],
        "serve_p50_ms": st["p50_ms"],
        "serve_p99_ms": st["p99_ms"],
        "serve_rate_hz": rate,
        "serve_closed_loop_s": busy,
        "serve_wall_s": wall,
        "serve_updates_per_s": work / wall,
        "serve_peak_resident_gb": eng.peak_resident / 1e9,
        "serve_max_qubits": int(max(r.n for r in reqs)),
    }


def cpu_baseline(w):
    """qandle's complex64 torch semantics restated (oracle/qandle_c64.py, structured
    mode: the dense reference needs 2 x 8 TiB of cached matrices per gate at 20
    qubits) on a bounded sample: the full circuit on as many of the 64 samples as fit
    in 10-30 s on this host (8 samples)."""
    import torch

    from oracle import qandle_c64

    threads = min(CPU_THREADS, os.cpu_count() or 1)
    torch.set_num_threads(threads)

    def run(Bs):
        ops = w.torch_ops()
        ops[0] = (ops[0][0], ops[0][1], ops[0][2], ops[0][3][:Bs])
        t0 = time.perf_counter()
        qandle_c64.structured_run(w.n, ops, Bs)
        return time.perf_counter() - t0

    run(1)  # warm-up
    Bs = 8  # 8 of 64 samples: 54 MB of states, well outside the caches; about 10-15 s here
    dt = run(Bs)
    updates = w.gate_ops * Bs * 2 ** w.n
    return {
        "value": updates / dt,
        "unit": "amplitude-gate updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full {w.gate_ops}-gate circuit on {Bs} of the {w.B} samples, torch CPU complex64 "
                  f"(structured restatement of qandle's gate semantics), {dt:.1f} s",
        "cpu": _cpu_model(),
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
