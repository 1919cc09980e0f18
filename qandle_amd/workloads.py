"""Seeded synthetic circuit families used by the benchmark, the tests and the
serving loop's request generator.

Every family is built from qandle's own layers with explicit angles and
``remapping=None`` (the default ``qw_map.tanh`` remapping is not pinned by any
reference file).  ``Workload.oracle_ops()`` gives the same circuit in the
plain-tuple format of the CPU oracle.
"""
from __future__ import annotations

import dataclasses
import math
import typing

import numpy as np
import torch

from . import embeddings, measurements, operators


@dataclasses.dataclass
class Workload:
    name: str
    n: int
    B: int
    layers: list                  # unbuilt qandle_amd layers (embedding ... measurement)
    inputs: typing.Dict[str, torch.Tensor]
    gates: list                   # (kind, qubit | (c, t), angle) in program order, without the embedding
    embed: tuple                  # (rot, qubits, input name)
    loss_w: typing.Optional[torch.Tensor] = None

    @property
    def gate_ops(self) -> int:
        """Gate applications per forward (embedded qubits count as one gate each)."""
        return len(self.embed[1]) + len(self.gates)

    def amplitude_updates(self) -> int:
        return self.gate_ops * self.B * 2 ** self.n

    def oracle_ops(self):
        rot, qubits, name = self.embed
        ops = [("embed", rot, list(qubits), self.inputs[name].double().numpy())]
        for g in self.gates:
            if g[0] == "cnot":
                ops.append(("cnot", g[1][0], g[1][1]))
            else:
                ops.append((g[0], g[1], float(g[2])))
        return ops

    def torch_ops(self):
        """Oracle tuples with float32 torch angles (for the complex64 restatement)."""
        rot, qubits, name = self.embed
        ops = [("embed", rot, list(qubits), self.inputs[name])]
        for g in self.gates:
            if g[0] == "cnot":
                ops.append(("cnot", g[1][0], g[1][1]))
            else:
                ops.append((g[0], g[1], torch.tensor(g[2], dtype=torch.float32)))
        return ops


def _uniform(gen, shape, lo, hi):
    return torch.rand(shape, generator=gen, dtype=torch.float32) * (hi - lo) + lo


def _assemble(name, n, B, x, body, measure=True, embed_rot="x"):
    layers = [embeddings.AngleEmbedding(name="x", qubits=list(range(n)), rotation="r" + embed_rot)]
    ctor = {"rx": operators.RX, "ry": operators.RY, "rz": operators.RZ}
    for g in body:
        if g[0] == "cnot":
            layers.append(operators.CNOT(g[1][0], g[1][1]))
        else:
            layers.append(ctor[g[0]](qubit=g[1], theta=torch.tensor(g[2]), remapping=None))
    if measure:
        layers.append(measurements.MeasureProbability())
    return Workload(name, n, B, layers, {"x": x}, body, (embed_rot, tuple(range(n)), "x"))


def embed_ry_ring(n=8, B=32, depth=4, seed=0):
    """RX AngleEmbedding, depth x (RY per qubit + CNOT ring q -> q+1 mod n), MeasureProbability."""
    gen = torch.Generator().manual_seed(seed)
    x = _uniform(gen, (B, n), 0.0, math.pi)
    body = []
    for _ in range(depth):
        for q in range(n):
            body.append(("ry", q, float(_uniform(gen, (), 0.0, 2 * math.pi))))
        for q in range(n):
            body.append(("cnot", (q, (q + 1) % n)))
    return _assemble(f"embed_ry_ring_n{n}_b{B}_d{depth}", n, B, x, body)


def rxryrz_ladder(n=20, B=64, depth=10, seed=1, loss=False):
    """RX AngleEmbedding, depth x (RX, RY, RZ on every qubit + CNOT ladder q -> q+1), MeasureProbability."""
    gen = torch.Generator().manual_seed(seed)
    x = _uniform(gen, (B, n), 0.0, math.pi)
    body = []
    for _ in range(depth):
        for q in range(n):
            for k in ("rx", "ry", "rz"):
                body.append((k, q, float(_uniform(gen, (), 0.0, 2 * math.pi))))
        for q in range(n - 1):
            body.append(("cnot", (q, q + 1)))
    w = _assemble(f"rxryrz_ladder_n{n}_b{B}_d{depth}", n, B, x, body)
    if loss:
        w.loss_w = torch.randn((B, n), generator=gen, dtype=torch.float32)
    return w


def random_uneven(num_circuits=2048, n_lo=4, n_hi=18, d_lo=2, d_hi=24, p_cnot=0.3, seed=3, build_layers=True):
    """Circuits of uneven width and depth: per layer one random gate from {RX, RY, RZ}
    per qubit (angles U[0, 2pi)) and a CNOT on each adjacent pair q -> q+1 with
    probability p; batch 1; RX AngleEmbedding with inputs U[0, pi)."""
    rng = np.random.default_rng(seed)
    kinds = ("rx", "ry", "rz")
    out = []
    for c in range(num_circuits):
        n = int(rng.integers(n_lo, n_hi + 1))
        depth = int(rng.integers(d_lo, d_hi + 1))
        body = []
        for _ in range(depth):
            ks = rng.integers(0, 3, n)
            th = rng.uniform(0.0, 2 * math.pi, n).astype(np.float32)
            for q in range(n):
                body.append((kinds[ks[q]], q, float(th[q])))
            coin = rng.random(max(n - 1, 0))
            for q in range(n - 1):
                if coin[q] < p_cnot:
                    body.append(("cnot", (q, q + 1)))
        x = torch.from_numpy(rng.uniform(0.0, math.pi, (1, n)).astype(np.float32))
        if build_layers:
            out.append(_assemble(f"uneven_{c}_n{n}_d{depth}", n, 1, x, body))
        else:
            out.append(Workload(f"uneven_{c}_n{n}_d{depth}", n, 1, [], {"x": x}, body, ("x", tuple(range(n)), "x")))
    return out


def hardware_efficient(n, B, depth, gen):
    """RY/RZ on every qubit + CNOT ring, ``depth`` times (the serving request ansatz)."""
    x = _uniform(gen, (B, n), 0.0, math.pi)
    body = []
    for _ in range(depth):
        for q in range(n):
            body.append(("ry", q, float(_uniform(gen, (), 0.0, 2 * math.pi))))
            body.append(("rz", q, float(_uniform(gen, (), 0.0, 2 * math.pi))))
        for q in range(n):
            body.append(("cnot", (q, (q + 1) % n)))
    return _assemble(f"hea_n{n}_b{B}_d{depth}", n, B, x, body)
