"""Helpers mirrored from ``src/qandle/utils.py``."""
from __future__ import annotations

import typing

import torch


def parse_rot(rot: str):
    """``src/qandle/utils.py:26-37``: x/rx/y/ry/z/rz in any case -> operator class."""
    from . import operators as op

    r = rot.lower().replace("r", "")
    if r == "x":
        return op.RX
    if r == "y":
        return op.RY
    if r == "z":
        return op.RZ
    raise ValueError(f"Unknown rotation {rot}")


def do_not_implement(*protected, reason="") -> type:
    """Metaclass factory forbidding overrides of ``protected`` names (``src/qandle/utils.py:40-51``)."""

    class LimitedClass(type):
        def __new__(cls, name, bases, attrs):
            for attribute in attrs:
                if attribute in protected:
                    if reason:
                        raise AttributeError(reason)
                    raise AttributeError(f'Overiding of attribute "{attribute}" not allowed')
            return super().__new__(cls, name, bases, attrs)

    return LimitedClass


def get_matrix_transforms(num_qubits: int, in_circuit: typing.List[int]):
    """State <-> sub-matrix views for a sub-circuit (``src/qandle/utils.py:54-70``).

    Returns two callables: ``to_matrix`` maps (batch, 2^n) to
    (batch * 2^(n-k), 2^k) with the sub-circuit qubits (in the given order) as
    the minor axis; ``to_state`` is its inverse.  The engine never materialises
    these views on the hot path (sub-circuits are lowered onto parent bit
    positions); they exist for API compatibility.
    """
    n = num_qubits
    inside = list(in_circuit)
    outside = [q for q in range(n) if q not in inside]
    perm = outside + inside

    def to_matrix(x: torch.Tensor) -> torch.Tensor:
        b = x.shape[0]
        v = x.reshape((b,) + (2,) * n).permute((0,) + tuple(p + 1 for p in perm))
        return v.reshape(b * 2 ** len(outside), 2 ** len(inside))

    inv = [perm.index(q) for q in range(n)]

    def to_state(x: torch.Tensor) -> torch.Tensor:
        b = x.shape[0] // 2 ** len(outside)
        v = x.reshape((b,) + (2,) * n).permute((0,) + tuple(p + 1 for p in inv))
        return v.reshape(b, 2 ** n)

    return to_matrix, to_state


def reduce_dot(*args):
    """Matrix product of a series of (batched) matrices (``src/qandle/utils.py:6-23``)."""
    if len(args) == 1:
        return args[0]
    matrix = args[0]
    if matrix.dim() == 2:
        matrix = matrix.unsqueeze(0)
    for sub in args[1:]:
        if sub.dim() == 2:
            sm = sub.unsqueeze(0)
        elif sub.dim() == 3:
            sm = sub
        else:
            raise ValueError(f"Invalid shape {sub.shape}")
        matrix = torch.matmul(sm, matrix)
    return matrix
