"""Operators of the drop-in API (``src/qandle/operators.py``).

Same class names, constructor contracts, ``build()`` protocol, ``named``
dispatch flag, ``state_dict`` keys (``...theta``) and registry
(``BUILT_CLASS_RELATION`` with its ``.T`` inverse) as the reference.  What
changed is underneath: no operator holds a 2^n x 2^n gate-matrix cache.  A
rotation is applied by the HIP kernel ``qs_rot`` from two scalars
(cos(phi/2), sin(phi/2)); CNOT is the exact permutation ``qs_cnot``.  Inside a
``Circuit`` the operators are not called one by one at all: the circuit lowers
them to a program for the fused engine.
"""
from __future__ import annotations

import abc
import typing

import torch

from . import config
from . import errors
from . import qasm
from . import qw_map
from . import utils
from ._engine import plain

__all__ = [
    "Operator",
    "RX",
    "RY",
    "RZ",
    "CNOT",
    "U",
    "CustomGate",
    "Invert",
    "Controlled",
    "CZ",
    "SWAP",
    "Reset",
    "BUILT_CLASS_RELATION",
]


class AbstractNoForward(abc.ABCMeta, utils.do_not_implement("forward", "backward")):
    """Forbids ``forward``/``backward`` on unbuilt operators (``src/qandle/operators.py:31-32``)."""


def qubit_bit(num_qubits: int, qubit: int) -> int:
    """Amplitude-index bit position of qandle qubit ``qubit``.

    qandle orders the Kronecker product with qubit 0 first
    (``src/qandle/operators.py:284-286``), so qubit q is bit n-1-q
    (``:549``).  This is the only place the engine converts.
    """
    if not 0 <= qubit < num_qubits:
        raise ValueError(f"qubit {qubit} out of range for {num_qubits} qubits")
    return num_qubits - 1 - qubit


class Operator(abc.ABC):
    """Everything that can be applied to a state (``src/qandle/operators.py:35-49``)."""

    named = False

    @abc.abstractmethod
    def __str__(self) -> str:
        """String representation."""

    def __repr__(self) -> str:
        return self.__str__()

    def to_qasm(self) -> qasm.QasmRepresentation:
        """The OpenQASM statement(s) of the operator (``src/qandle/operators.py:47-49``)."""
        raise NotImplementedError(f"{type(self).__name__} has no OpenQASM form")


class UnbuiltOperator(Operator, abc.ABC):
    """Operators that have not been built yet (``src/qandle/operators.py:52-57``)."""

    @abc.abstractmethod
    def build(self, num_qubits, **kwargs) -> "BuiltOperator":
        """Converts the operator to a torch.nn.Module for ``num_qubits`` qubits."""


class BuiltOperator(Operator, torch.nn.Module, abc.ABC):
    """Built operators (``src/qandle/operators.py:60-69``)."""

    @abc.abstractmethod
    def forward(self, state: torch.Tensor) -> torch.Tensor:
        """Applies the operator to the state."""

    @abc.abstractmethod
    def to_matrix(self, **kwargs) -> torch.Tensor:
        """Matrix M with ``state @ M == forward(state)``."""

    def _act(self, st: torch.Tensor, cmask: int = 0, inverse: bool = False, **kwargs) -> torch.Tensor:
        """The operator on a (B, 2^n) state, only where every bit position in
        ``cmask`` is 1, optionally inverted (what Controlled and Invert build on)."""
        raise NotImplementedError(f"{type(self).__name__} cannot be controlled or inverted")

    def _act_matrix(self, st: torch.Tensor, cmask: int = 0, inverse: bool = False, **kwargs) -> torch.Tensor:
        """The operator that ``to_matrix`` describes (equal to ``_act`` except
        where the reference's ``to_matrix`` differs from its ``forward``)."""
        return self._act(st, cmask, inverse, **kwargs)


# ----------------------------------------------------------------------------
# batch handling shared by the single-operator forwards (S5 of the reference:
# an unbatched state with a batched named input yields a batched output,
# ``src/qandle/operators.py:290-298``)
# ----------------------------------------------------------------------------

def _as_batch(state: torch.Tensor, num_qubits: int):
    if state.shape[-1] != 2 ** num_qubits:
        raise RuntimeError(
            f"state has {state.shape[-1]} amplitudes, operator expects {2 ** num_qubits}"
        )
    if state.dim() == 1:
        return state.unsqueeze(0), True
    if state.dim() == 2:
        return state, False
    raise RuntimeError("qandle_amd operators take (2^n,) or (B, 2^n) states")


def _forward(op: "BuiltOperator", state: torch.Tensor, **kwargs) -> torch.Tensor:
    from ._engine.lib import require_device

    require_device(state, "state")
    st, unbatched = _as_batch(state, op.num_qubits)
    out = op._act(st, **kwargs)
    return out.squeeze(0) if unbatched and out.shape[0] == 1 else out


def _dense_1q(num_qubits: int, qubit: int, matrix: torch.Tensor) -> torch.Tensor:
    """kron over the qubits, ``matrix`` at ``qubit`` (qubit 0 first), as in
    ``src/qandle/operators.py:101-104``."""
    m = torch.eye(1, dtype=torch.complex64)
    for i in range(num_qubits):
        m = torch.kron(m, matrix.to(torch.complex64) if i == qubit else torch.eye(2, dtype=torch.complex64))
    return m


class UnbuiltParametrizedOperator(UnbuiltOperator, metaclass=AbstractNoForward):
    """Parametrised rotation before build (``src/qandle/operators.py:150-212``)."""

    def __init__(
        self,
        qubit: int,
        theta: typing.Union[float, torch.Tensor, None] = None,
        name: typing.Union[str, None] = None,
        **kwargs,
    ):
        assert isinstance(qubit, int), "qubit must be an integer"
        assert qubit >= 0, "qubit must be >= 0"
        if isinstance(theta, (float, int)):
            theta = torch.tensor(theta, requires_grad=True, dtype=torch.float)
        remapping = kwargs.get("remapping", config.DEFAULT_MAPPING)
        if remapping is None:
            remapping = qw_map.none
        self.qubit = qubit
        self.name = name
        self.named = name is not None
        self.theta = theta
        self.remapping = remapping

    def __str__(self) -> str:
        if self.name is None:
            if self.theta is None:
                return f"{self.__class__.__name__}{self.qubit}"
            return f"{self.__class__.__name__}{self.qubit} ({self.remapping(self.theta).item():.2f})"
        return f"{self.__class__.__name__}{self.qubit} ({self.name})"

    def to_qasm(self) -> qasm.QasmRepresentation:
        """``rx(<angle>) q[i];``, or the input name for a named gate (``src/qandle/operators.py:186-204``)."""
        gate = self.__class__.__name__.lower()
        if self.named:
            return qasm.QasmRepresentation(gate_str=gate, qubit=self.qubit, qasm3_inputs=self.name)
        if self.theta is None:
            raise errors.UnbuiltGateError(
                "This gate has no parameter. Set parameter or build or set name before converting to OpenQASM2."
            )
        return qasm.QasmRepresentation(gate_str=gate, qubit=self.qubit, gate_value=self.remapping(self.theta).item())

    def build(self, num_qubits, **kwargs) -> "BuiltParametrizedOperator":
        return BUILT_CLASS_RELATION[self.__class__](
            qubit=self.qubit,
            initialtheta=self.theta,
            name=self.name,
            remapping=self.remapping,
            num_qubits=num_qubits,
        )


class BuiltParametrizedOperator(BuiltOperator, abc.ABC):
    """Built rotation (``src/qandle/operators.py:215-305``) without the matrix cache.

    ``theta`` is an ``nn.Parameter`` initialised exactly as the reference does
    (``torch.rand(1)`` when none is given, consumed at build time).
    """

    KIND = -1  # 0 = X, 1 = Y, 2 = Z (the Pauli generator)

    def __init__(
        self,
        qubit: int,
        remapping: typing.Callable,
        num_qubits: int,
        initialtheta: typing.Union[torch.Tensor, None] = None,
        name: typing.Union[str, None] = None,
    ):
        super().__init__()
        self.qubit = qubit
        if initialtheta is None:
            initialtheta = torch.rand(1)
        self.name = name
        self.named = name is not None
        self.theta = torch.nn.Parameter(initialtheta, requires_grad=True)
        self.remapping = remapping
        self.num_qubits = num_qubits
        self.unbuilt_class = BUILT_CLASS_RELATION.T[self.__class__]
        self.bit = qubit_bit(num_qubits, qubit)

    def __str__(self) -> str:
        base = f"{self.unbuilt_class.__name__}_{self.qubit}"
        if self.named:
            return f"{base} ({self.name})"
        return f"{base} ({self.remapping(self.theta).item():.2f})"

    def to_qasm(self) -> qasm.QasmRepresentation:
        """As the unbuilt gate's (``src/qandle/operators.py:250-262``)."""
        gate = self.unbuilt_class.__name__.lower()
        if self.named:
            return qasm.QasmRepresentation(gate_str=gate, qubit=self.qubit, qasm3_inputs=self.name)
        return qasm.QasmRepresentation(gate_str=gate, qubit=self.qubit, gate_value=self.remapping(self.theta).item())

    def angle(self, **kwargs) -> torch.Tensor:
        """phi: remap(theta) for an unnamed gate, kwargs[name] (no remap) for a named one
        (``src/qandle/operators.py:265-271``)."""
        if self.named:
            return kwargs[self.name]
        return self.remapping(self.theta)

    def forward(self, state: torch.Tensor, **kwargs) -> torch.Tensor:
        from ._engine.lib import require_device

        require_device(state, "state")
        st, unbatched = _as_batch(state, self.num_qubits)
        out = self._act(st, **kwargs)
        return out.squeeze(0) if unbatched and out.shape[0] == 1 else out

    def _act(self, st: torch.Tensor, cmask: int = 0, inverse: bool = False, **kwargs) -> torch.Tensor:
        phi = self.angle(**kwargs)
        if not isinstance(phi, torch.Tensor):
            phi = torch.tensor(phi, dtype=torch.float32)
        phi = phi.to(device=st.device, dtype=torch.float32)
        if phi.dim() > 1:
            raise RuntimeError("named rotation inputs must be 0-d or (B,)")
        if phi.numel() == 1 and phi.dim() == 1 and not self.named:
            phi = phi.reshape(())
        if phi.dim() == 1:
            if st.shape[0] == 1 and phi.shape[0] != 1:
                st = st.expand(phi.shape[0], -1)
            elif st.shape[0] != phi.shape[0]:
                raise RuntimeError(
                    f"batched input of size {phi.shape[0]} does not match state batch {st.shape[0]}"
                )
        if inverse:  # R_P(phi)^-1 = R_P(-phi)
            phi = -phi
        if cmask:
            return plain.ControlledRotFn.apply(st, phi, self.num_qubits, self.bit, self.KIND, cmask)
        return plain.RotFn.apply(st, phi, self.num_qubits, self.bit, self.KIND)

    def to_matrix(self, **kwargs) -> torch.Tensor:
        """Dense ``M = U^T`` expanded to 2^n x 2^n (a debugging aid; never on the hot path)."""
        phi = self.angle(**kwargs)
        if not isinstance(phi, torch.Tensor):
            phi = torch.tensor(phi, dtype=torch.float32)
        t = phi / 2
        c, s = torch.cos(t), torch.sin(t)
        z = torch.zeros_like(c)
        if self.KIND == 0:
            u = torch.stack([torch.complex(c, z), torch.complex(z, -s), torch.complex(z, -s), torch.complex(c, z)], -1)
        elif self.KIND == 1:
            u = torch.stack([torch.complex(c, z), torch.complex(-s, z), torch.complex(s, z), torch.complex(c, z)], -1)
        else:
            u = torch.stack([torch.complex(c, -s), torch.complex(z, z), torch.complex(z, z), torch.complex(c, s)], -1)
        u = u.reshape(u.shape[:-1] + (2, 2))
        m = torch.ones((1, 1), dtype=torch.complex64)
        eye = torch.eye(2, dtype=torch.complex64)
        for w in range(self.num_qubits):
            m = torch.kron(m, u if w == self.qubit else eye) if u.dim() == 2 else _bkron(m, u if w == self.qubit else eye)
        return m.transpose(-1, -2)


def _bkron(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    if a.dim() == 2:
        a = a.unsqueeze(0)
    if b.dim() == 2:
        b = b.unsqueeze(0)
    out = a[:, :, None, :, None] * b[:, None, :, None, :]
    return out.reshape(out.shape[0], a.shape[1] * b.shape[1], a.shape[2] * b.shape[2])


class RX(UnbuiltParametrizedOperator):
    """Parametrised RX gate, exp(-i phi X / 2) (``src/qandle/operators.py:308-325``)."""


class RY(UnbuiltParametrizedOperator):
    """Parametrised RY gate, exp(-i phi Y / 2) (``src/qandle/operators.py:328-345``)."""


class RZ(UnbuiltParametrizedOperator):
    """Parametrised RZ gate, exp(-i phi Z / 2) (``src/qandle/operators.py:348-365``)."""


class BuiltRX(BuiltParametrizedOperator):
    KIND = 0


class BuiltRY(BuiltParametrizedOperator):
    KIND = 1


class BuiltRZ(BuiltParametrizedOperator):
    KIND = 2


class CNOT(UnbuiltOperator):
    """CNOT gate (``src/qandle/operators.py:398-413``)."""

    def __init__(self, control: int, target: int):
        assert control != target, "Control and target must be different"
        self.c = control
        self.t = target

    def __str__(self) -> str:
        return f"CNOT {self.c}|{self.t}"

    def to_qasm(self) -> qasm.QasmRepresentation:
        return qasm.QasmRepresentation(gate_str=f"cx q[{self.c}], q[{self.t}]")

    def build(self, num_qubits, **kwargs) -> "BuiltCNOT":
        return BuiltCNOT(control=self.c, target=self.t, num_qubits=num_qubits)


class BuiltCNOT(BuiltOperator):
    """CNOT as an exact permutation (``src/qandle/operators.py:536-567``), no 4^n matrix."""

    def __init__(self, control: int, target: int, num_qubits: int):
        super().__init__()
        self.c = control
        self.t = target
        self.num_qubits = num_qubits
        self.cbit = qubit_bit(num_qubits, control)
        self.tbit = qubit_bit(num_qubits, target)

    def forward(self, state: torch.Tensor) -> torch.Tensor:
        return _forward(self, state)

    def _act(self, st: torch.Tensor, cmask: int = 0, inverse: bool = False, **kwargs) -> torch.Tensor:
        if cmask & (1 << self.tbit):
            raise ValueError("a control qubit of Controlled is also the CNOT's target")
        if cmask:  # CNOT is its own inverse
            return plain.McxFn.apply(st, self.num_qubits, self.tbit, cmask | (1 << self.cbit))
        return plain.CnotFn.apply(st, self.num_qubits, self.cbit, self.tbit)

    def __str__(self) -> str:
        return CNOT(self.c, self.t).__str__()

    def to_qasm(self) -> qasm.QasmRepresentation:
        return CNOT(self.c, self.t).to_qasm()

    def to_matrix(self, **kwargs) -> torch.Tensor:
        N = 2 ** self.num_qubits
        idx = torch.arange(N)
        src = torch.where((idx & (1 << self.cbit)) != 0, idx ^ (1 << self.tbit), idx)
        M = torch.zeros(N, N, dtype=torch.complex64)
        M[idx, src] = 1
        return M


class U(UnbuiltOperator):
    """Custom one-qubit gate before build (``src/qandle/operators.py:72-84``)."""

    def __init__(self, qubit: int, matrix: torch.Tensor):
        self.qubit = qubit
        self.matrix = matrix

    def __str__(self) -> str:
        return f"U_{self.qubit}"

    def to_qasm(self) -> qasm.QasmRepresentation:
        return qasm.QasmRepresentation(gate_str="U", qubit=self.qubit)

    def build(self, num_qubits, **kwargs) -> "BuiltU":
        return BuiltU(qubit=self.qubit, matrix=self.matrix, num_qubits=num_qubits)


class BuiltU(BuiltOperator):
    """Custom one-qubit gate (``src/qandle/operators.py:87-129``).

    The reference multiplies the state by the 2^n x 2^n kron of ``matrix``
    (``state @ M``); here the kernel ``qs_unitary_1q`` applies the 2x2 block
    to each amplitude pair, in the column convention U = matrix^T.  The
    dense ``matrix`` of the reference is built only on request."""

    def __init__(self, qubit: int, matrix: torch.Tensor, num_qubits: int, self_description: str = "U"):
        super().__init__()
        self.qubit = qubit
        self.num_qubits = num_qubits
        self.description = self_description
        self.original_matrix = matrix
        self.bit = qubit_bit(num_qubits, qubit)
        m = torch.as_tensor(matrix).detach().to(torch.complex128).cpu()
        if m.shape != (2, 2):
            raise ValueError("U takes a 2x2 matrix")
        self.register_buffer("_u", _pack_u(m.T), persistent=False)
        self.register_buffer("_u_inv", _pack_u(torch.linalg.inv(m).T), persistent=False)

    @property
    def matrix(self) -> torch.Tensor:
        return _dense_1q(self.num_qubits, self.qubit, torch.as_tensor(self.original_matrix)).contiguous()

    def __str__(self) -> str:
        return f"{self.description}_{self.qubit}"

    def to_qasm(self) -> qasm.QasmRepresentation:
        """A gate definition holding the four entries to two decimals (``src/qandle/operators.py:109-124``;
        the doubled braces are the reference's text)."""
        m = torch.as_tensor(self.original_matrix)
        entries = ", ".join(format(m[i, j].item(), ".2f") for i, j in ((0, 0), (0, 1), (1, 0), (1, 1)))
        q = self.qubit
        return qasm.QasmRepresentation(gate_str=f"gate {self.description}_{q} q[{q}] {{{{ U({entries}) }}}}", qubit=q)

    def forward(self, state: torch.Tensor) -> torch.Tensor:
        return _forward(self, state)

    def _act(self, st: torch.Tensor, cmask: int = 0, inverse: bool = False, **kwargs) -> torch.Tensor:
        if cmask & (1 << self.bit):
            raise ValueError("a control qubit of Controlled is also the gate's qubit")
        u = self._u_inv if inverse else self._u
        return plain.UnitaryFn.apply(st, u.to(st.device), self.num_qubits, self.bit, cmask)

    def to_matrix(self, **kwargs) -> torch.Tensor:
        return self.matrix


def _pack_u(u: torch.Tensor) -> torch.Tensor:
    """8 float32 (re, im of u00, u01, u10, u11) of a 2x2 complex matrix."""
    return torch.view_as_real(u.to(torch.complex64).contiguous()).reshape(8).clone()


CustomGate = BuiltU


class Invert(UnbuiltOperator):
    """Inverse of a target operator (``src/qandle/operators.py:416-431``)."""

    def __init__(self, target: Operator):
        self.t = target

    def __str__(self) -> str:
        return f"{self.t}^-1"

    def to_qasm(self) -> qasm.QasmRepresentation:
        return qasm.QasmRepresentation(gate_str=f"{self.t}^-1")

    def build(self, num_qubits, **kwargs) -> "BuiltInvert":
        return BuiltInvert(target=self.t, num_qubits=num_qubits)


class BuiltInvert(BuiltOperator):
    """Applies the inverse of what the target's ``to_matrix`` describes
    (``src/qandle/operators.py:434-456``): rotations with the negated angle,
    U with the inverse matrix, CNOT/CZ/SWAP as themselves."""

    def __init__(self, target: Operator, num_qubits: int):
        super().__init__()
        if hasattr(target, "build"):
            import warnings

            warnings.warn(
                "Building target operator in Invert. If the target operator is parameterized, this might lead to "
                "different parameter initialization. Please build the target operator before inverting it."
            )
            target = target.build(num_qubits)
        self.target = target
        self.num_qubits = num_qubits

    def forward(self, state: torch.Tensor) -> torch.Tensor:
        return _forward(self, state)

    def _act(self, st: torch.Tensor, cmask: int = 0, inverse: bool = False, **kwargs) -> torch.Tensor:
        # the reference inverts target.to_matrix() called without keyword inputs
        return self.target._act_matrix(st, cmask, not inverse)

    def __str__(self) -> str:
        return Invert(self.target).__str__()

    def to_qasm(self) -> qasm.QasmRepresentation:
        return Invert(self.target).to_qasm()

    def to_matrix(self, **kwargs) -> torch.Tensor:
        return torch.linalg.inv(self.target.to_matrix())


class Controlled(UnbuiltOperator):
    """Target operator applied where a control qubit is 1 (``src/qandle/operators.py:459-475``)."""

    def __init__(self, control: int, target: Operator):
        self.c = control
        self.t = target

    def __str__(self) -> str:
        return f"Controlled {self.c}|{self.t}"

    def to_qasm(self) -> qasm.QasmRepresentation:
        return qasm.QasmRepresentation(gate_str=f"controlled q[{self.c}], {self.t}")

    def build(self, num_qubits, **kwargs) -> "BuiltControlled":
        return BuiltControlled(control=self.c, target=self.t, num_qubits=num_qubits)


class BuiltControlled(BuiltOperator):
    """Controlled target (``src/qandle/operators.py:478-515``).

    The reference applies ``target.to_matrix(**kwargs)`` to the control = 1
    half of the state; here the target's kernel runs with the control bit in
    its control mask.  As in the reference, ``to_matrix`` (and so Invert or an
    outer Controlled of this operator) describes the target alone."""

    def __init__(self, control: int, target: Operator, num_qubits: int):
        super().__init__()
        self.c = control
        if hasattr(target, "build"):
            target = target.build(num_qubits)
        self.t = target
        self.named = target.named
        self.num_qubits = num_qubits
        self.cbit = qubit_bit(num_qubits, control)

    def forward(self, state: torch.Tensor, **kwargs) -> torch.Tensor:
        return _forward(self, state, **kwargs)

    def _act(self, st: torch.Tensor, cmask: int = 0, inverse: bool = False, **kwargs) -> torch.Tensor:
        return self.t._act_matrix(st, cmask | (1 << self.cbit), inverse, **kwargs)

    def _act_matrix(self, st: torch.Tensor, cmask: int = 0, inverse: bool = False, **kwargs) -> torch.Tensor:
        return self.t._act_matrix(st, cmask, inverse, **kwargs)

    def __str__(self) -> str:
        return Controlled(self.c, self.t).__str__()

    def to_qasm(self) -> qasm.QasmRepresentation:
        return Controlled(self.c, self.t).to_qasm()

    def to_matrix(self, **kwargs) -> torch.Tensor:
        return self.t.to_matrix()


class CZ(UnbuiltOperator):
    """CZ gate (``src/qandle/operators.py:518-533``)."""

    def __init__(self, control: int, target: int):
        assert control != target, "Control and target must be different"
        self.c = control
        self.t = target

    def __str__(self) -> str:
        return f"CZ {self.c}|{self.t}"

    def to_qasm(self) -> qasm.QasmRepresentation:
        return qasm.QasmRepresentation(gate_str=f"cz q[{self.c}], q[{self.t}]")

    def build(self, num_qubits, **kwargs) -> "BuiltCZ":
        return BuiltCZ(control=self.c, target=self.t, num_qubits=num_qubits)


class BuiltCZ(BuiltOperator):
    """CZ as a sign flip of the amplitudes with both bits set
    (``src/qandle/operators.py:570-600``), no 4^n matrix."""

    def __init__(self, control: int, target: int, num_qubits: int):
        super().__init__()
        self.c = control
        self.t = target
        self.num_qubits = num_qubits
        self.mask = (1 << qubit_bit(num_qubits, control)) | (1 << qubit_bit(num_qubits, target))

    def forward(self, state: torch.Tensor) -> torch.Tensor:
        return _forward(self, state)

    def _act(self, st: torch.Tensor, cmask: int = 0, inverse: bool = False, **kwargs) -> torch.Tensor:
        return plain.PhaseFlipFn.apply(st, self.num_qubits, self.mask | cmask)  # its own inverse

    def __str__(self) -> str:
        return CZ(self.c, self.t).__str__()

    def to_qasm(self) -> qasm.QasmRepresentation:
        return CZ(self.c, self.t).to_qasm()

    def to_matrix(self, **kwargs) -> torch.Tensor:
        idx = torch.arange(2 ** self.num_qubits)
        diag = torch.ones(2 ** self.num_qubits, dtype=torch.complex64)
        diag[(idx & self.mask) == self.mask] = -1
        return torch.diag(diag)


class Reset(UnbuiltOperator):
    """Reset of one qubit to |0> keeping each pair's norm (``src/qandle/operators.py:635-650``)."""

    def __init__(self, qubit: int):
        self.qubit = qubit

    def __str__(self) -> str:
        return f"Reset {self.qubit}"

    def to_qasm(self) -> qasm.QasmRepresentation:
        return qasm.QasmRepresentation(gate_str="reset", qubit=self.qubit)

    def build(self, num_qubits, **kwargs) -> "BuiltReset":
        return BUILT_CLASS_RELATION[self.__class__](qubit=self.qubit, num_qubits=num_qubits)


class BuiltReset(BuiltOperator):
    """Reset (``src/qandle/operators.py:603-632``): for every pair of amplitudes
    that differ only in this qubit, a0 <- a0 * |(a0, a1)| / (|a0| + 1e-7) and
    a1 <- 0 (kernel ``qs_reset``).  Not unitary: no matrix, so it can be neither
    inverted nor controlled."""

    def __init__(self, qubit: int, num_qubits: int):
        super().__init__()
        self.qubit = qubit
        self.num_qubits = num_qubits
        self.bit = qubit_bit(num_qubits, qubit)

    def forward(self, state: torch.Tensor) -> torch.Tensor:
        from ._engine.lib import require_device

        require_device(state, "state")
        st, unbatched = _as_batch(state, self.num_qubits)
        out = plain.ResetFn.apply(st, self.num_qubits, self.bit)
        return out.squeeze(0) if unbatched else out

    def _act_matrix(self, st: torch.Tensor, cmask: int = 0, inverse: bool = False, **kwargs) -> torch.Tensor:
        raise ValueError("Reset gate does not have a matrix representation.")

    def __str__(self) -> str:
        return Reset(self.qubit).__str__()

    def to_qasm(self) -> qasm.QasmRepresentation:
        return Reset(self.qubit).to_qasm()

    def to_matrix(self, **kwargs) -> torch.Tensor:
        raise ValueError("Reset gate does not have a matrix representation.")


class SWAP(UnbuiltOperator):
    """SWAP gate (``src/qandle/operators.py:686-698``)."""

    def __init__(self, a: int, b: int):
        self.a = a
        self.b = b

    def __str__(self) -> str:
        return f"Swap {self.a}|{self.b}"

    def to_qasm(self) -> qasm.QasmRepresentation:
        return qasm.QasmRepresentation(gate_str=f"swap q[{self.a}], q[{self.b}]")

    def build(self, num_qubits, **kwargs) -> "BuiltSWAP":
        return BuiltSWAP(a=self.a, b=self.b, num_qubits=num_qubits)


class BuiltSWAP(BuiltOperator):
    """SWAP as an exact permutation (``src/qandle/operators.py:653-683``); the
    same qubit twice is the identity, as the reference's matrix is."""

    def __init__(self, a: int, b: int, num_qubits: int):
        super().__init__()
        self.a = a
        self.b = b
        self.num_qubits = num_qubits
        self.abit = qubit_bit(num_qubits, a)
        self.bbit = qubit_bit(num_qubits, b)

    def forward(self, state: torch.Tensor) -> torch.Tensor:
        return _forward(self, state)

    def _act(self, st: torch.Tensor, cmask: int = 0, inverse: bool = False, **kwargs) -> torch.Tensor:
        if self.abit == self.bbit:
            return st.clone()
        if cmask & ((1 << self.abit) | (1 << self.bbit)):
            raise ValueError("a control qubit of Controlled is also a SWAP qubit")
        return plain.SwapFn.apply(st, self.num_qubits, self.abit, self.bbit, cmask)  # its own inverse

    def __str__(self) -> str:
        return f"SWAP {self.a}|{self.b}"

    def to_qasm(self) -> qasm.QasmRepresentation:
        return SWAP(self.a, self.b).to_qasm()

    def to_matrix(self, **kwargs) -> torch.Tensor:
        N = 2 ** self.num_qubits
        idx = torch.arange(N)
        differ = ((idx >> self.abit) & 1) != ((idx >> self.bbit) & 1)
        src = torch.where(differ, idx ^ ((1 << self.abit) | (1 << self.bbit)), idx)
        M = torch.zeros(N, N, dtype=torch.complex64)
        M[idx, src] = 1
        return M


class rdict(dict):
    """Reversible dict (``src/qandle/operators.py:701-706``)."""

    @property
    def T(self):
        return {v: k for k, v in self.items()}


BUILT_CLASS_RELATION = rdict(
    {
        UnbuiltOperator: BuiltOperator,
        UnbuiltParametrizedOperator: BuiltParametrizedOperator,
        RX: BuiltRX,
        RY: BuiltRY,
        RZ: BuiltRZ,
        CNOT: BuiltCNOT,
        Reset: BuiltReset,
        U: BuiltU,
        SWAP: BuiltSWAP,
        CZ: BuiltCZ,
        Invert: BuiltInvert,
        Controlled: BuiltControlled,
    }
)
