"""Writes the golden vectors in this folder from the FP64 oracle (oracle/fp64.py).

    python tests/golden/make_golden.py

Each .npz holds inputs and expected outputs only.  The oracle itself is
pinned by tests/test_oracle.py (closed forms, expm, brute force, finite
differences); these files freeze its outputs so any later change to the
oracle or to a kernel shows up as a diff against fixed numbers.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import fp64  # noqa: E402


def kat_rotations():
    """RX/RY/RZ at the reference tests' angles on seeded states, n = 1..5, batch 13."""
    rng = np.random.default_rng(42)
    out = {}
    for v, tag in ((0.543321, "v1"), (-2.3456, "v2")):
        for n in range(1, 6):
            psi = rng.uniform(size=(13, 2 ** n)) + 1j * rng.uniform(size=(13, 2 ** n))
            psi = (psi / np.linalg.norm(psi, axis=1, keepdims=True)).astype(np.complex64)
            out[f"{tag}_n{n}_in"] = psi
            for q in range(n):
                for k in "xyz":
                    U = fp64.rot_matrix(k, float(np.float32(v)))
                    out[f"{tag}_n{n}_q{q}_r{k}"] = fp64.apply_1q(psi.astype(np.complex128), n, q, U).astype(np.complex64)
    return out


def c1_family():
    """8 qubits, batch 32: RX embedding, 4 x (RY per qubit + CNOT ring), MeasureProbability."""
    rng = np.random.default_rng(0)
    n, B = 8, 32
    X = rng.uniform(0, np.pi, size=(B, n)).astype(np.float32)
    th = rng.uniform(0, 2 * np.pi, size=(4, n)).astype(np.float32)
    ops = [("embed", "x", list(range(n)), X.astype(np.float64))]
    for l in range(4):
        for q in range(n):
            ops.append(("ry", q, float(th[l, q])))
        for q in range(n):
            ops.append(("cnot", q, (q + 1) % n))
    psi = fp64.simulate(n, ops)
    w = rng.normal(size=(B, n))
    g = fp64.adjoint_grads(n, ops, w)
    gth = np.array([a for a, op in zip(g[1:], ops[1:]) if op[0] != "cnot"])
    return {"X": X, "theta": th, "w": w, "psi": psi.astype(np.complex64), "P": fp64.measure_prob(psi, n),
            "grad_X": g[0], "grad_theta": gth}


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "kat_rotations.npz"), **kat_rotations())
    np.savez_compressed(os.path.join(HERE, "c1_family.npz"), **c1_family())
    print("written")
