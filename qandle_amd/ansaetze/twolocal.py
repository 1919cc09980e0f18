"""TwoLocal (``src/qandle/ansaetze/twolocal.py:13-123``)."""
from __future__ import annotations

import copy
import typing

import torch

from .. import config
from .. import operators as op
from .ansatz import BuiltAnsatz, UnbuiltAnsatz

__all__ = ["TwoLocal"]


def column_cnot_chain(state: torch.Tensor, qubits: typing.List[int], num_qubits: int) -> torch.Tensor:
    """The reference's ``cnot_matrices @ state`` (``src/qandle/ansaetze/twolocal.py:106``,
    ``specialunitary.py:106``): the product C_0 C_1 ... of the chain's CNOT
    matrices applied from the left.  On a (2^n,) state that is the chain in
    reverse order, last CNOT first (each CNOT matrix is a symmetric permutation).
    On a batch it is a matrix product over the batch axis, defined only when
    the batch has 2^n rows: it permutes the rows, which the engine reproduces."""
    if len(qubits) < 2:
        raise IndexError("list index out of range")  # the reference's cnots[0] on an empty chain
    if state.dim() == 1:
        for ci in reversed(range(len(qubits) - 1)):
            state = op.CNOT(qubits[ci], qubits[ci + 1]).build(num_qubits)(state)
        return state
    if state.shape[0] != state.shape[-1]:
        raise RuntimeError(
            f"mat1 and mat2 shapes cannot be multiplied ({state.shape[-1]}x{state.shape[-1]} and "
            f"{state.shape[0]}x{state.shape[-1]})"
        )
    # rows: out[i] = state[f_0(f_1(...f_k(i)))] with f_j the j-th CNOT's index map
    idx = torch.arange(state.shape[0])
    for ci in range(len(qubits) - 1):
        cb = num_qubits - 1 - qubits[ci]
        tb = num_qubits - 1 - qubits[ci + 1]
        idx = torch.where((idx >> cb) & 1 == 1, idx ^ (1 << tb), idx)
    return state.index_select(0, idx.to(state.device))


class TwoLocal(UnbuiltAnsatz):
    """Alternating RY layers and a linear CNOT chain (``:13-60``)."""

    def __init__(self, qubits: typing.Union[typing.List[int], None] = None, depth: int = 1,
                 remapping: typing.Union[typing.Callable, None] = config.DEFAULT_MAPPING, q_params=None):
        super().__init__()
        self.depth = depth
        self.remapping = remapping
        self.q_params = q_params
        self.qubits = qubits

    def build(self, *args, **kwargs) -> BuiltAnsatz:
        return TwoLocalBuilt(num_qubits=kwargs["num_qubits"], qubits=self.qubits, depth=self.depth,
                             remapping=self.remapping, q_params=self.q_params)

    def __str__(self) -> str:
        return "TL"

    def decompose(self) -> typing.List[op.Operator]:
        if self.qubits is None:
            raise ValueError(
                "It is not specified on which qubits to apply the ansatz. Build the circuit to specify or pass "
                "them explicitly as a list"
            )
        res = []
        qp = self.q_params if self.q_params is not None else torch.rand(self.depth, len(self.qubits))
        for d in range(self.depth):
            for wi in range(len(self.qubits)):
                res.append(op.RY(qubit=self.qubits[wi], theta=qp[d, wi]))
            for wi in range(len(self.qubits) - 1):
                res.append(op.CNOT(self.qubits[wi], self.qubits[wi + 1]))
        return res


class TwoLocalBuilt(BuiltAnsatz):
    """Built TwoLocal (``:63-123``).  As in the reference, the RY layers cover
    every circuit qubit with the default remapping, ``q_params`` of another
    shape than (depth, len(qubits)) is redrawn, and the CNOT chain is applied
    as ``cnot_matrices @ state`` (see :func:`column_cnot_chain`)."""

    def __init__(self, num_qubits: int, qubits: typing.Union[typing.List[int], None], depth: int = 1,
                 remapping: typing.Union[typing.Callable, None] = config.DEFAULT_MAPPING, q_params=None):
        super().__init__()
        self.num_qubits = num_qubits
        self.depth = depth
        self.remapping = remapping
        self.qubits = qubits or list(range(num_qubits))
        if q_params is None or q_params.shape != (depth, len(self.qubits)):
            q_params = torch.rand(depth, num_qubits)
        layers = [[op.RY(qubit=w, theta=q_params[d, w]).build(num_qubits) for w in range(num_qubits)]
                  for d in range(depth)]
        self.mods = torch.nn.ModuleList([torch.nn.Sequential(*lay) for lay in layers])

    def forward(self, state: torch.Tensor):
        for d in range(self.depth):
            state = self.mods[d](state)
            state = column_cnot_chain(state, self.qubits, self.num_qubits)
        return state

    def __str__(self) -> str:
        return "TL"

    def decompose(self) -> typing.List[op.Operator]:
        outp = []
        for mod in self.mods:
            for seq in mod:
                outp.append(copy.deepcopy(seq))
            for ci in range(len(self.qubits) - 1):
                outp.append(op.CNOT(self.qubits[ci], self.qubits[ci + 1]))
        return outp
