"""Restatement of qandle's complex64 torch path on the CPU (TEST INFRASTRUCTURE ONLY).

Two modes, same op-tuple format as :mod:`oracle.fp64` but with float32 torch
angle tensors:

``dense``
    What the reference itself executes.  Every rotation gets the two cached
    2^n x 2^n "partial matrices" built by an n-fold Kronecker product
    (``src/qandle/operators.py:277-287``), stored transposed
    (``:236-238``), recombined each call as ``_a*a_op(t) + _b*b_op(t)`` with
    t = phi/2 (``:265-275``) and applied as ``state @ M`` (``:289-298``; a
    batched angle gives a (B, N, N) stack and a bmm, ``:294-297``).  CNOT is the
    dense permutation ``_M`` (``:547-558``).  AngleEmbedding is the einsum +
    chained-matmul product applied to |0..0> (``src/qandle/embeddings.py:132-164``)
    and MeasureProbability the repeat/mask/reduce of
    ``src/qandle/measurements.py:109-123``.  Feasible to about n = 10.
``structured``
    Same float32 half-angle sin/cos, the same gate entries, applied as 2x2
    complex64 updates on a (B, 2^q, 2, 2^(n-1-q)) view, so that it runs at the
    workload sizes where the dense cache cannot exist (a 20-qubit cache is
    8 TiB per matrix).

This module is also the ``cpu_baseline`` port timed by ``bench.py``: it is
what qandle's torch code computes, restated, because the reference itself
cannot be imported (its hard dependency ``qw_map`` is not installed).
"""
from __future__ import annotations

import torch

_I = torch.tensor(1j, dtype=torch.cfloat)


# ---------------------------------------------------------------- dense mode

def _hydrated(n: int, q: int, special=None) -> torch.Tensor:
    """n-fold kron with ``special`` at qubit q (``src/qandle/operators.py:277-287``)."""
    if special is None:
        special = torch.eye(2)
    elif not isinstance(special, torch.Tensor):
        special = torch.tensor(special)
    m = torch.eye(1)
    for w in range(n):
        m = torch.kron(m, special if w == q else torch.eye(2))
    return m.to(torch.cfloat)


def _builder(kind: str, n: int, q: int):
    """(a, b, a_op, b_op) of ``src/qandle/operators.py:368-395``."""
    if kind == "x":
        return -_hydrated(n, q, ((0, 1), (1, 0))) * _I, _hydrated(n, q), torch.sin, torch.cos
    if kind == "y":
        return _hydrated(n, q, ((0, -1), (1, 0))), _hydrated(n, q), torch.sin, torch.cos
    if kind == "z":
        return (
            _hydrated(n, q, ((1, 0), (0, 0))),
            _hydrated(n, q, ((0, 0), (0, 1))),
            lambda t: torch.exp(-_I * t),
            lambda t: torch.exp(_I * t),
        )
    raise ValueError(kind)


class DenseGate:
    """One rotation with its gate-matrix cache (built once, like the reference)."""

    def __init__(self, kind: str, n: int, q: int):
        a, b, self.a_op, self.b_op = _builder(kind, n, q)
        self._a = a.T.contiguous()
        self._b = b.T.contiguous()

    def __call__(self, state: torch.Tensor, phi: torch.Tensor) -> torch.Tensor:
        t = phi / 2
        if t.dim() == 1:
            t = t.unsqueeze(-1).unsqueeze(-1)
        mat = self._a * self.a_op(t) + self._b * self.b_op(t)
        if mat.dim() == 2:
            return state @ mat
        if state.dim() == 1:
            state = state.unsqueeze(0)
        return (state.unsqueeze(1) @ mat).squeeze(1)


def dense_cnot_matrix(n: int, c: int, t: int) -> torch.Tensor:
    M = torch.zeros(2 ** n, 2 ** n) * 0j
    c2, t2 = n - c - 1, n - t - 1
    for i in range(2 ** n):
        if i & (1 << c2):
            M[i, i ^ (1 << t2)] = 1
        else:
            M[i, i] = 1
    return M


class DenseEmbedding:
    """AngleEmbedding as the reference evaluates it (``src/qandle/embeddings.py:132-164``)."""

    def __init__(self, rot: str, n: int, qubits):
        gates = [_builder(rot, n, q) for q in qubits]
        self._a = torch.stack([g[0] for g in gates], 0).contiguous()
        self._b = torch.stack([g[1] for g in gates], 0).contiguous()
        self.a_op, self.b_op = gates[0][2], gates[0][3]
        self._z = torch.zeros(2 ** n, dtype=torch.cfloat)
        self._z[0] = 1

    def __call__(self, X: torch.Tensor) -> torch.Tensor:
        a = torch.einsum("dab,...d->...dab", self._a, self.a_op(X / 2))
        b = torch.einsum("dab,...d->...dab", self._b, self.b_op(X / 2))
        m = a + b
        res = m[..., 0, :, :]
        for i in range(1, m.shape[-3]):
            res = res @ m[..., i, :, :]
        return res @ self._z


def dense_measure_prob(state: torch.Tensor, n: int) -> torch.Tensor:
    """``src/qandle/measurements.py:109-123`` including its squeeze quirks."""
    ind = []
    for w in range(n):
        ind.append((torch.arange(2 ** n) // (2 ** w)) % 2 == 0)
    indices0 = torch.stack(ind, 0).flip(dims=(0,))
    rep = state.repeat(n, 1, 1)
    sums = rep * indices0[:, None, :]
    return sums.abs().pow(2).sum(dim=-1).squeeze().transpose(-1, 0)


class DenseCircuit:
    """Builds every cache up front (the reference's build step), then runs."""

    def __init__(self, n: int, ops):
        self.n = n
        self.steps = []
        for op in ops:
            if op[0] == "embed":
                self.steps.append(("embed", DenseEmbedding(op[1], n, op[2])))
            elif op[0] in ("rx", "ry", "rz"):
                self.steps.append(("rot", DenseGate(op[0][1], n, op[1])))
            elif op[0] == "cnot":
                self.steps.append(("cnot", dense_cnot_matrix(n, op[1], op[2])))
            else:
                raise ValueError(op[0])

    def __call__(self, ops, state=None, measure=True):
        n = self.n
        if state is None:
            state = torch.zeros(2 ** n, dtype=torch.complex64)
            state[0] = 1
        for (kind, obj), op in zip(self.steps, ops):
            if kind == "embed":
                state = obj(_as_f32(op[3]))
            elif kind == "rot":
                state = obj(state, _as_f32(op[2]))
            else:
                state = state @ obj
        return dense_measure_prob(state, n) if measure else state


# ------------------------------------------------------------ structured mode

def _as_f32(a) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.to(torch.float32)
    return torch.as_tensor(a, dtype=torch.float32)


def _coeffs(kind: str, phi: torch.Tensor):
    """Entries (u00, u01, u10, u11) as complex64, from float32 sin/cos of phi/2."""
    t = phi / 2
    if kind == "z":
        e_m = torch.exp(-_I * t)
        e_p = torch.exp(_I * t)
        z = torch.zeros_like(e_m)
        return e_m, z, z, e_p
    s, c = torch.sin(t), torch.cos(t)
    cc = torch.complex(c, torch.zeros_like(c))
    if kind == "x":
        ms = torch.complex(torch.zeros_like(s), -s)
        return cc, ms, ms, cc
    ss = torch.complex(s, torch.zeros_like(s))
    return cc, -ss, ss, cc


def structured_rot(state: torch.Tensor, n: int, q: int, kind: str, phi: torch.Tensor) -> torch.Tensor:
    B = state.shape[0]
    v = state.reshape(B, 2 ** q, 2, 2 ** (n - 1 - q))
    u00, u01, u10, u11 = _coeffs(kind, phi)
    if phi.dim() == 1:
        u00, u01, u10, u11 = (u.reshape(-1, 1, 1) for u in (u00, u01, u10, u11))
    a0, a1 = v[:, :, 0, :], v[:, :, 1, :]
    out = torch.stack((u00 * a0 + u01 * a1, u10 * a0 + u11 * a1), dim=2)
    return out.reshape(B, 2 ** n)


def structured_cnot(state: torch.Tensor, n: int, c: int, t: int) -> torch.Tensor:
    idx = torch.arange(2 ** n)
    cm, tm = 1 << (n - 1 - c), 1 << (n - 1 - t)
    src = torch.where((idx & cm) != 0, idx ^ tm, idx)
    return state[:, src]


def structured_embed(X: torch.Tensor, n: int, qubits, rot: str) -> torch.Tensor:
    if X.dim() == 1:
        X = X.unsqueeze(0)
    B = X.shape[0]
    state = torch.zeros(B, 2 ** n, dtype=torch.complex64)
    state[:, 0] = 1
    for j, q in enumerate(qubits):
        state = structured_rot(state, n, q, rot, X[:, j])
    return state


def structured_measure_prob(state: torch.Tensor, n: int) -> torch.Tensor:
    B = state.shape[0]
    p = state.abs().pow(2)
    cols = []
    for q in range(n):
        cols.append(p.reshape(B, 2 ** q, 2, 2 ** (n - 1 - q))[:, :, 0, :].sum(dim=(1, 2)))
    return torch.stack(cols, 1)


def structured_run(n: int, ops, B: int, state=None, measure=True):
    if state is None:
        state = torch.zeros(B, 2 ** n, dtype=torch.complex64)
        state[:, 0] = 1
    for op in ops:
        if op[0] == "embed":
            state = structured_embed(_as_f32(op[3]), n, op[2], op[1])
        elif op[0] in ("rx", "ry", "rz"):
            phi = _as_f32(op[2])
            state = structured_rot(state, n, op[1], op[0][1], phi)
        elif op[0] == "cnot":
            state = structured_cnot(state, n, op[1], op[2])
        else:
            raise ValueError(op[0])
    return structured_measure_prob(state, n) if measure else state
