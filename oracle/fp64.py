"""complex128 state-vector oracle for qandle's gate path (TEST INFRASTRUCTURE ONLY).

Conventions restated from the reference (file:line into gstenzel/qandle v0.1.8):

* state: (B, 2^n) amplitudes; qubit q is amplitude-index bit n-1-q
  (``src/qandle/operators.py:284-286`` kron order; ``:549`` ``c2 = n-c-1``).
* gates act as psi' = U psi in column convention; the reference stores
  M = U^T and computes ``state @ M`` (``src/qandle/operators.py:236-238, 292``).
* RX = [[c, -is], [-is, c]], RY = [[c, -s], [s, c]], RZ = diag(e^{-it}, e^{it})
  with t = phi/2, c = cos t, s = sin t (``src/qandle/operators.py:368-395``,
  combined at ``:272-274``).
* CNOT: amplitudes i and i ^ (1 << (n-1-t)) swap when bit n-1-c of i is set
  (``src/qandle/operators.py:547-555``).
* AngleEmbedding: the incoming state is ignored and the output is
  (x)_q R(x_q)|0..0> over the embedded qubits (``src/qandle/embeddings.py:148-164``).
* MeasureProbability: P[b, q] = sum_{i : bit n-1-q of i is 0} |psi_b[i]|^2
  (``src/qandle/measurements.py:113-123``).

A circuit is a list of plain tuples so that the oracle shares no code with the
product:

    ("embed", rot, qubits, X)   rot in "xyz", X: (B, d) or (d,) angles
    ("rx" | "ry" | "rz", q, phi)  phi: float (shared) or (B,) array (per sample)
    ("cnot", c, t)
    ("u", q, M)                  custom gate, M the reference's 2x2 ``matrix``
    ("cz", c, t) | ("swap", a, b) | ("reset", q)
    ("ctrl", c, op)              op (one of the above gate tuples) where qubit c is 1
    ("inv", op)                  inverse of a gate tuple

The remaining operators (``src/qandle/operators.py:72-147, 416-699``):

* U(q, M): ``state @ kron(.., M at q, ..)`` (``:87-126``), i.e. psi' = M^T psi.
* Invert(T): ``state @ inv(T.to_matrix())`` (``:445-447``).
* Controlled(c, T): T's matrix on the c = 1 half, the c = 0 half unchanged
  (``:488-506``); its own ``to_matrix`` is T's (``:514-515``).
* CZ: -1 on the amplitudes with bits c and t set (``:581-591``).
* SWAP: exchange of bits a and b (``:662-674``).
* Reset(q): per amplitude pair (a0: q = 0, a1: q = 1),
  a0' = a0 |(a0, a1)| / (|a0| + 1e-7), a1' = 0 (``:612-623``).
"""
from __future__ import annotations

import numpy as np

PAULI = {
    "x": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


def bit(n: int, q: int) -> int:
    """Amplitude-index bit of qandle qubit q (``src/qandle/operators.py:549``)."""
    return n - 1 - q


def rot_matrix(kind: str, phi) -> np.ndarray:
    """2x2 column-convention matrices, closed forms of ``src/qandle/operators.py:368-395``.

    ``phi`` may be a scalar or an array; the result has shape phi.shape + (2, 2).
    """
    phi = np.asarray(phi, dtype=np.float64)
    t = phi / 2.0
    c, s = np.cos(t), np.sin(t)
    out = np.zeros(phi.shape + (2, 2), dtype=np.complex128)
    if kind == "x":
        out[..., 0, 0] = c
        out[..., 1, 1] = c
        out[..., 0, 1] = -1j * s
        out[..., 1, 0] = -1j * s
    elif kind == "y":
        out[..., 0, 0] = c
        out[..., 1, 1] = c
        out[..., 0, 1] = -s
        out[..., 1, 0] = s
    elif kind == "z":
        out[..., 0, 0] = np.exp(-1j * t)
        out[..., 1, 1] = np.exp(1j * t)
    else:
        raise ValueError(kind)
    return out


def apply_1q(psi: np.ndarray, n: int, q: int, U: np.ndarray) -> np.ndarray:
    """psi' = U psi on qubit q.  U is (2, 2) or (B, 2, 2) (per-sample)."""
    B = psi.shape[0]
    v = psi.reshape(B, 2 ** q, 2, 2 ** (n - 1 - q))
    if U.ndim == 2:
        out = np.einsum("ij,bajc->baic", U, v)
    else:
        out = np.einsum("bij,bajc->baic", U, v)
    return out.reshape(B, 2 ** n)


def apply_cnot(psi: np.ndarray, n: int, c: int, t: int) -> np.ndarray:
    idx = np.arange(2 ** n)
    cm, tm = 1 << bit(n, c), 1 << bit(n, t)
    src = np.where(idx & cm, idx ^ tm, idx)
    return psi[:, src]


def apply_u(psi: np.ndarray, n: int, q: int, M: np.ndarray, ctrl=()) -> np.ndarray:
    """U(q, M) (``src/qandle/operators.py:87-126``): psi' = M^T psi on qubit q,
    only where every qubit in ``ctrl`` is 1 (``:488-506``)."""
    out = apply_1q(psi, n, q, np.swapaxes(np.asarray(M, dtype=np.complex128), -1, -2))
    if not ctrl:
        return out
    on = _ctrl_mask(n, ctrl)
    return np.where(on[None, :], out, psi)


def _ctrl_mask(n: int, ctrl) -> np.ndarray:
    idx = np.arange(2 ** n)
    on = np.ones(2 ** n, dtype=bool)
    for c in ctrl:
        on &= ((idx >> bit(n, c)) & 1) == 1
    return on


def apply_cz(psi: np.ndarray, n: int, c: int, t: int, ctrl=()) -> np.ndarray:
    """CZ (``src/qandle/operators.py:581-591``): -1 where bits c and t are set."""
    on = _ctrl_mask(n, tuple(ctrl) + (c, t))
    return np.where(on[None, :], -psi, psi)


def apply_swap(psi: np.ndarray, n: int, a: int, b: int, ctrl=()) -> np.ndarray:
    """SWAP (``src/qandle/operators.py:662-674``): out[i] = psi[i with bits a, b exchanged]."""
    idx = np.arange(2 ** n)
    ma, mb = 1 << bit(n, a), 1 << bit(n, b)
    differ = ((idx & ma) != 0) != ((idx & mb) != 0)
    src = np.where(differ & _ctrl_mask(n, ctrl), idx ^ (ma | mb), idx)
    return psi[:, src]


def apply_reset(psi: np.ndarray, n: int, q: int) -> np.ndarray:
    """Reset (``src/qandle/operators.py:612-623``), per pair of amplitudes that
    differ only in qubit q: a0' = a0 * |(a0, a1)| / (|a0| + 1e-7), a1' = 0."""
    B = psi.shape[0]
    v = psi.reshape(B, 2 ** q, 2, 2 ** (n - 1 - q))
    a0, a1 = v[:, :, 0, :], v[:, :, 1, :]
    norm = np.sqrt(np.abs(a0) ** 2 + np.abs(a1) ** 2)
    out = np.zeros_like(v)
    out[:, :, 0, :] = a0 * norm / (np.abs(a0) + 1e-7)
    return out.reshape(B, 2 ** n)


def amplitude_embed(X: np.ndarray, n: int, pad=None, normalize: bool = False) -> np.ndarray:
    """AmplitudeEmbedding (``src/qandle/embeddings.py:50-61``): the features
    padded with ``pad`` to 2^n entries (``F.pad``; a longer row is cropped),
    divided by max(||row||, 1e-12) when ``normalize`` (``F.normalize``), as
    real amplitudes."""
    X = np.atleast_2d(np.asarray(X, dtype=np.float64))
    N = 2 ** n
    if X.shape[1] >= N:
        V = X[:, :N].copy()
    else:
        if pad is None:
            raise ValueError("features do not fill the register and no padding value is given")
        V = np.concatenate([X, np.full((X.shape[0], N - X.shape[1]), float(pad))], axis=1)
    if normalize:
        V = V / np.maximum(np.linalg.norm(V, axis=1, keepdims=True), 1e-12)
    return V.astype(np.complex128)


def joint_prob(psi: np.ndarray) -> np.ndarray:
    """MeasureJointProbability (``src/qandle/measurements.py:78-79``): |psi|^2."""
    return np.abs(psi) ** 2


def embed(X: np.ndarray, n: int, qubits, rot: str) -> np.ndarray:
    """(x)_j R(x_j)|0> over ``qubits``; unembedded qubits stay |0>."""
    X = np.asarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X[None, :]
    B = X.shape[0]
    psi = np.zeros((B, 2 ** n), dtype=np.complex128)
    psi[:, 0] = 1.0
    for j, q in enumerate(qubits):
        psi = apply_1q(psi, n, q, rot_matrix(rot, X[:, j]))
    return psi


def measure_prob(psi: np.ndarray, n: int) -> np.ndarray:
    """(B, n) float64, P(qubit q = 0)."""
    B = psi.shape[0]
    p = np.abs(psi) ** 2
    out = np.empty((B, n))
    for q in range(n):
        v = p.reshape(B, 2 ** q, 2, 2 ** (n - 1 - q))
        out[:, q] = v[:, :, 0, :].sum(axis=(1, 2))
    return out


def _batch_of(ops, psi0):
    B = 1 if psi0 is None else psi0.shape[0]
    for op in ops:
        if op[0] == "embed":
            X = np.asarray(op[3])
            if X.ndim == 2:
                B = max(B, X.shape[0])
        elif op[0] in ("rx", "ry", "rz"):
            a = np.asarray(op[2])
            if a.ndim == 1:
                B = max(B, a.shape[0])
    return B


def simulate(n: int, ops, psi0: np.ndarray | None = None, record: bool = False):
    """Run ``ops``; returns psi (B, 2^n) [and the list of states after each op]."""
    B = _batch_of(ops, psi0)
    if psi0 is None:
        psi = np.zeros((B, 2 ** n), dtype=np.complex128)
        psi[:, 0] = 1.0
    else:
        psi = np.broadcast_to(np.asarray(psi0, dtype=np.complex128), (B, 2 ** n)).copy()
    states = []
    for op in ops:
        kind = op[0]
        if kind == "embed":
            _, rot, qubits, X = op
            X = np.asarray(X, dtype=np.float64)
            if X.ndim == 1:
                X = np.broadcast_to(X, (B, X.shape[0]))
            psi = embed(X, n, qubits, rot)
        elif kind in ("rx", "ry", "rz"):
            U = rot_matrix(kind[1], op[2])
            if U.ndim == 3 and U.shape[0] != B:
                U = np.broadcast_to(U, (B, 2, 2))
            psi = apply_1q(psi, n, op[1], U)
        elif kind == "cnot":
            psi = apply_cnot(psi, n, op[1], op[2])
        elif kind in ("u", "cz", "swap", "reset", "ctrl", "inv"):
            psi = apply_gate(psi, n, op)
        else:
            raise ValueError(kind)
        if record:
            states.append(psi.copy())
    return (psi, states) if record else psi


def gate_matrix_1q(op):
    """(qubit, reference-convention 2x2 matrix M) of a one-qubit gate tuple
    (``state @ M`` form), for ``ctrl`` and ``inv``."""
    kind = op[0]
    if kind in ("rx", "ry", "rz"):
        return op[1], np.swapaxes(rot_matrix(kind[1], op[2]), -1, -2)
    if kind == "u":
        return op[1], np.asarray(op[2], dtype=np.complex128)
    return None


def apply_gate(psi: np.ndarray, n: int, op, ctrl=()) -> np.ndarray:
    """One gate tuple, with extra control qubits ``ctrl``."""
    kind = op[0]
    if kind in ("rx", "ry", "rz", "u"):
        q, M = gate_matrix_1q(op)
        return apply_u(psi, n, q, M, ctrl)
    if kind == "cnot":
        return apply_u(psi, n, op[2], np.array([[0, 1], [1, 0]]), tuple(ctrl) + (op[1],))
    if kind == "cz":
        return apply_cz(psi, n, op[1], op[2], ctrl)
    if kind == "swap":
        return psi.copy() if op[1] == op[2] else apply_swap(psi, n, op[1], op[2], ctrl)
    if kind == "reset":
        if ctrl:
            raise ValueError("Reset gate does not have a matrix representation.")
        return apply_reset(psi, n, op[1])
    if kind == "ctrl":
        # the target's matrix (an inner Controlled contributes only its own target)
        inner = op[2]
        while inner[0] == "ctrl":
            inner = inner[2]
        return apply_gate(psi, n, inner, tuple(ctrl) + (op[1],))
    if kind == "inv":
        return apply_gate(psi, n, inverse_of(op[1]), ctrl)
    raise ValueError(kind)


def inverse_of(op):
    """Gate tuple of inv(op.to_matrix()) (``src/qandle/operators.py:445-447``)."""
    kind = op[0]
    if kind in ("rx", "ry", "rz"):
        return (kind, op[1], -np.asarray(op[2]))
    if kind == "u":
        return ("u", op[1], np.linalg.inv(np.asarray(op[2], dtype=np.complex128)))
    if kind in ("cnot", "cz", "swap"):
        return op
    if kind == "ctrl":  # to_matrix of a Controlled is its target's
        inner = op[2]
        while inner[0] == "ctrl":
            inner = inner[2]
        return inverse_of(inner)
    if kind == "inv":
        return op[1]
    raise ValueError("Reset gate does not have a matrix representation.")


# ----------------------------------------------------------------------------
# dense-unitary route (independent of apply_1q's reshape): explicit Kronecker
# products with wire 0 as the most significant factor.
# ----------------------------------------------------------------------------

def kron_1q(n: int, q: int, U: np.ndarray) -> np.ndarray:
    m = np.eye(1, dtype=np.complex128)
    for w in range(n):
        m = np.kron(m, U if w == q else np.eye(2))
    return m


def cnot_matrix(n: int, c: int, t: int) -> np.ndarray:
    """Brute-force permutation matrix, column convention (psi' = M psi)."""
    N = 2 ** n
    M = np.zeros((N, N), dtype=np.complex128)
    for i in range(N):
        bits = [(i >> (n - 1 - w)) & 1 for w in range(n)]
        if bits[c]:
            bits[t] ^= 1
        j = sum(b << (n - 1 - w) for w, b in enumerate(bits))
        M[j, i] = 1.0
    return M


def dense_unitary(n: int, ops) -> np.ndarray:
    """Full 2^n x 2^n unitary for gate-only circuits with shared angles."""
    Umat = np.eye(2 ** n, dtype=np.complex128)
    for op in ops:
        if op[0] in ("rx", "ry", "rz"):
            G = kron_1q(n, op[1], rot_matrix(op[0][1], float(op[2])))
        elif op[0] == "cnot":
            G = cnot_matrix(n, op[1], op[2])
        else:
            raise ValueError(op[0])
        Umat = G @ Umat
    return Umat


# ----------------------------------------------------------------------------
# gradients of L = sum_{b,q} w[b,q] * P[b,q]
# ----------------------------------------------------------------------------

def loss(n: int, ops, w: np.ndarray, psi0=None) -> float:
    return float((measure_prob(simulate(n, ops, psi0), n) * w).sum())


def adjoint_grads(n: int, ops, w: np.ndarray, psi0=None):
    """FP64 adjoint sweep.  Returns a list aligned with ``ops``:

    * rotation op -> dL/dphi (scalar for a shared angle, (B,) for per-sample)
    * embed op    -> dL/dX with the shape of X
    * cnot        -> None

    Seed: lambda = 2 psi * sum_q w_q [bit_q(i) = 0]  (torch's dL/dRe + i dL/dIm).
    Per gate g = exp(-i phi P / 2):  dL/dphi = 1/2 Im <lambda_g | P | psi_g>
    at the gate output; then psi, lambda <- U^dag psi, U^dag lambda.
    """
    psi = simulate(n, ops, psi0)  # the sweep un-applies gates: no state is recorded
    B = psi.shape[0]
    w = np.broadcast_to(np.asarray(w, dtype=np.float64), (B, n))
    mask0 = np.zeros((n, 2 ** n))
    idx = np.arange(2 ** n)
    for q in range(n):
        mask0[q] = ((idx >> bit(n, q)) & 1) == 0
    lam = 2.0 * psi * (w @ mask0)
    grads = [None] * len(ops)
    for k in range(len(ops) - 1, -1, -1):
        op = ops[k]
        kind = op[0]
        if kind in ("rx", "ry", "rz"):
            P = PAULI[kind[1]]
            Ppsi = apply_1q(psi, n, op[1], P)
            g = 0.5 * np.imag(np.sum(np.conj(lam) * Ppsi, axis=1))
            grads[k] = g if np.asarray(op[2]).ndim == 1 else g.sum()
            U = rot_matrix(kind[1], op[2])
            Ud = np.conj(np.swapaxes(U, -1, -2))
            if Ud.ndim == 3 and Ud.shape[0] != B:
                Ud = np.broadcast_to(Ud, (B, 2, 2))
            psi = apply_1q(psi, n, op[1], Ud)
            lam = apply_1q(lam, n, op[1], Ud)
        elif kind == "cnot":
            psi = apply_cnot(psi, n, op[1], op[2])
            lam = apply_cnot(lam, n, op[1], op[2])
        elif kind == "embed":
            _, rot, qubits, X = op
            X = np.asarray(X, dtype=np.float64)
            P = PAULI[rot]
            gX = np.zeros((B, len(qubits)))
            for j, q in enumerate(qubits):
                Ppsi = apply_1q(psi, n, q, P)
                gX[:, j] = 0.5 * np.imag(np.sum(np.conj(lam) * Ppsi, axis=1))
            grads[k] = gX if X.ndim == 2 else gX.sum(0)
            # the embedding discards the incoming state: nothing flows further back
            break
    return grads


def with_angle(ops, k: int, value):
    """Copy of ``ops`` with op k's angle (or embedding matrix) replaced."""
    out = list(ops)
    op = out[k]
    if op[0] == "embed":
        out[k] = (op[0], op[1], op[2], value)
    else:
        out[k] = (op[0], op[1], value)
    return out


def finite_diff_grad(n: int, ops, w, k: int, h: float = 1e-6, psi0=None):
    """Central finite differences of L w.r.t. op k's angle(s)."""
    op = ops[k]
    base = np.array(op[3] if op[0] == "embed" else op[2], dtype=np.float64)
    g = np.zeros_like(base)
    for j in np.ndindex(base.shape):
        p, m = base.copy(), base.copy()
        p[j] += h
        m[j] -= h
        g[j] = (loss(n, with_angle(ops, k, p), w, psi0) - loss(n, with_angle(ops, k, m), w, psi0)) / (2 * h)
    return g


def param_shift_grad(n: int, ops, w, k: int, psi0=None):
    """Parameter shift for exp(-i phi P/2): dL/dphi = [L(phi+pi/2) - L(phi-pi/2)] / 2."""
    op = ops[k]
    base = np.array(op[3] if op[0] == "embed" else op[2], dtype=np.float64)
    g = np.zeros_like(base)
    s = np.pi / 2
    for j in np.ndindex(base.shape):
        p, m = base.copy(), base.copy()
        p[j] += s
        m[j] -= s
        g[j] = 0.5 * (loss(n, with_angle(ops, k, p), w, psi0) - loss(n, with_angle(ops, k, m), w, psi0))
    return g
