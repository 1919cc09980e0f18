"""Circuit executor (``src/qandle/qcircuit.py``).

``Circuit`` keeps the reference's constructor, ``forward(state=None, **kw)``,
``@``, ``split``, ``decompose`` and module hierarchy (so ``state_dict`` keys
such as ``circuit.layers.3.theta`` are unchanged).  ``UnsplittedCircuit.forward``
lowers its layers once into a program for the fused HIP engine (see
``qandle_amd/_engine/fused.py``); a layer list the program format does not
cover runs layer by layer through the operators' own kernels.
"""
from __future__ import annotations

import re
import typing
import warnings

import torch

from . import config
from . import drawer
from . import measurements
from . import operators
from . import qasm
from . import splitter

__all__ = ["Circuit", "UnsplittedCircuit", "SplittedCircuit", "Subcircuit"]


class Circuit(torch.nn.Module):
    class DummyUnbuiltCircuit:
        """Unbuilt layer list handed to the splitter (``src/qandle/qcircuit.py:14-27``)."""

        def __init__(self, num_qubits: int, layers: typing.List[operators.Operator]):
            self.num_qubits = num_qubits
            self.layers = []
            for layer in layers:
                if hasattr(layer, "decompose"):
                    self.layers.extend(layer.decompose())
                else:
                    self.layers.append(layer)

        def decompose(self):
            return self

    def __init__(self, layers, num_qubits=None, split_max_qubits=0, circuit=None):
        super().__init__()
        self.name = ""
        self.named = True
        if not hasattr(layers, "__iter__"):
            layers = [layers]
        layers = list(layers)
        self.num_qubits = num_qubits if num_qubits is not None else len(self._used_qubits(layers))
        if circuit is not None:
            self.circuit = circuit
        elif split_max_qubits > 0 and self.num_qubits > split_max_qubits:
            self.circuit = SplittedCircuit(
                num_qubits=self.num_qubits,
                subcircuitdict=splitter.split(
                    self.DummyUnbuiltCircuit(self.num_qubits, layers), max_qubits=split_max_qubits
                ),
            )
        else:
            self.circuit = UnsplittedCircuit(self.num_qubits, layers)

    def forward(self, state=None, **kwargs):
        return self.circuit.forward(state, **kwargs)

    def to_matrix(self, **kwargs):
        return self.circuit.to_matrix(**kwargs)

    def __matmul__(self, x):
        return self.circuit.forward(x)

    @staticmethod
    def _used_qubits(layers) -> typing.Set[int]:
        """Infer the qubit count (``src/qandle/qcircuit.py:70-106``)."""
        qubits = set()
        for layer in layers:
            if isinstance(layer, operators.CNOT):
                qubits.add(layer.c)
                qubits.add(layer.t)
            elif hasattr(layer, "qubit"):
                qubits.add(layer.qubit)
            elif hasattr(layer, "num_qubits"):
                qubits.update(range(layer.num_qubits))
            elif hasattr(layer, "qubits"):
                qubits.update(layer.qubits)
            elif isinstance(layer, (measurements.UnbuiltMeasurement, measurements.BuiltMeasurement)):
                pass
            else:
                raise ValueError(
                    f"Unknown layer type {type(layer)}, number of qubits could not be inferred. "
                    "Pass :code:`num_qubits` to the circuit."
                )
        if len(qubits) == 0:
            raise ValueError(
                "Number of qubits could not be inferred from layers. Please provide num_qubits to the circuit directly."
            )
        return qubits

    def split(self, max_qubits):
        if isinstance(self.circuit, SplittedCircuit):
            warnings.warn("Circuit already split, returning original circuit.")
            return self
        splitdict = splitter.split(self.circuit.decompose(), max_qubits=max_qubits)
        return Circuit(layers=[], num_qubits=self.num_qubits, circuit=SplittedCircuit(self.num_qubits, splitdict))

    def decompose(self):
        return Circuit(layers=[], num_qubits=self.num_qubits, circuit=self.circuit.decompose())

    def draw(self) -> str:
        """Text diagram of the circuit (``src/qandle/qcircuit.py:108-110``)."""
        return drawer.draw(self.circuit)

    def to_qasm(self):
        return self.circuit.to_qasm()

    def to_openqasm2(self) -> str:
        return qasm.convert_to_qasm(self, qasm_version=2, include_header=True)

    def to_openqasm3(self) -> str:
        return qasm.convert_to_qasm(self, qasm_version=3, include_header=True)


class UnsplittedCircuit(torch.nn.Module):
    """Layer list executed in order (``src/qandle/qcircuit.py:137-212``)."""

    def __init__(self, num_qubits: int, layers: list):
        super().__init__()
        self.num_qubits = num_qubits
        # default |0..0> (``:148-149``); the engine creates it on the device of
        # the call, so it is not a buffer here either
        self.state = torch.zeros(2 ** num_qubits, dtype=torch.complex64)
        self.state[0] = 1
        self.layers = torch.nn.ModuleList(self._build_layers(layers, num_qubits))

    @staticmethod
    def _build_layers(layers: list, num_qubits: int) -> typing.List[torch.nn.Module]:
        built = []
        for la in layers:
            built.append(la.build(num_qubits=num_qubits) if hasattr(la, "build") else la)
        return built

    def _device(self, state, kwargs):
        if state is not None:
            return state.device
        for v in kwargs.values():
            if isinstance(v, torch.Tensor):
                return v.device
        for p in self.parameters():
            return p.device
        return torch.device("cpu")

    def forward(self, state=None, **kwargs):
        if config.FUSED:
            from ._engine import fused

            prog = fused.lower_circuit(self)
            if prog is not None:
                return fused.run(prog, self, state, kwargs)
        return self.forward_layers(state, **kwargs)

    def forward_layers(self, state=None, **kwargs):
        """Operator-by-operator execution (the plain per-gate path)."""
        if state is None:
            dev = self._device(state, kwargs)
            state = self.state.to(dev)
        for mod in self.layers:
            state = mod(state, **kwargs) if mod.named else mod(state)
        return state

    def decompose(self) -> "UnsplittedCircuit":
        new_layers = []
        for layer in self.layers:
            if hasattr(layer, "decompose"):
                if isinstance(layer, Circuit):
                    new_layers.extend(layer.decompose().circuit.layers)
                else:
                    new_layers.extend(layer.decompose())
            else:
                new_layers.append(layer)
        return UnsplittedCircuit(self.num_qubits, new_layers)

    def to_qasm(self):
        """Statements of every layer that has an OpenQASM form, in order (``src/qandle/qcircuit.py:176-181``)."""
        reps = []
        for layer in self.layers:
            if hasattr(layer, "to_qasm"):
                reps.extend(layer.to_qasm())
        return reps

    def split(self, **kwargs):
        if self.num_qubits == 1:
            warnings.warn("Can't split circuit with only 1 qubit")
            return self
        return SplittedCircuit(self.num_qubits, splitter.split(self, **kwargs))

    def to_matrix(self, **kwargs):
        from . import utils

        return utils.reduce_dot(*[mod.to_matrix(**kwargs) for mod in self.layers])


class SplittedCircuit(UnsplittedCircuit):
    """Circuit made of sub-circuits (``src/qandle/qcircuit.py:215-258``).

    The engine lowers every sub-circuit gate back onto its parent bit position
    (the sub-circuit mapping is monotone, ``:240-244``), so a split circuit runs
    as one fused program with results equal to the unsplit circuit.
    """

    def __init__(self, num_qubits: int, subcircuitdict):
        torch.nn.Module.__init__(self)
        self.num_qubits = num_qubits
        self.state = torch.zeros(2 ** num_qubits, dtype=torch.complex64)
        self.state[0] = 1
        self._check_qubit_order(subcircuitdict)
        self.subcircuits = torch.nn.ModuleList(
            [Subcircuit(num_qubits, subcircuitdict[i].mapping, subcircuitdict[i].layers) for i in subcircuitdict]
        )

    @staticmethod
    def _check_qubit_order(subcircuitdict):
        for sc in subcircuitdict.values():
            keys_sorted = sorted(sc.mapping.keys())
            values = [sc.mapping[k] for k in keys_sorted]
            assert values == sorted(values), f"Mapping {sc.mapping} is not ordered."

    def forward_layers(self, inp=None, **kwargs):
        if inp is None:
            inp = self.state.to(self._device(inp, kwargs))
        for sc in self.subcircuits:
            inp = sc(inp, **kwargs)
        return inp

    def to_qasm(self):
        """One statement list per sub-circuit, on parent qubits (``src/qandle/qcircuit.py:253-255``)."""
        return [sc.to_qasm() for sc in self.subcircuits]

    def to_matrix(self, **kwargs):
        raise NotImplementedError("Can't get matrix representation of a splitted circuit.")


class Subcircuit(torch.nn.Module):
    """One sub-circuit of a split circuit (``src/qandle/qcircuit.py:261-295``)."""

    def __init__(self, num_qubits_parent: int, mapping: typing.Dict[int, int], layers: list):
        super().__init__()
        self.num_qubits_parent = num_qubits_parent
        self.mapping = mapping
        self.layers = torch.nn.ModuleList(UnsplittedCircuit._build_layers(layers, num_qubits=len(mapping)))
        self.inverse = {v: k for k, v in mapping.items()}

    def forward(self, x, **kwargs):
        """Applies each gate at its parent bit position (no permute-copy of the state)."""
        from ._engine import plain

        batched = x.dim() == 2
        st = x if batched else x.unsqueeze(0)
        n = self.num_qubits_parent
        for mod in self.layers:
            if isinstance(mod, operators.BuiltCNOT):
                st = plain.CnotFn.apply(st, n, n - 1 - self.inverse[mod.c], n - 1 - self.inverse[mod.t])
            elif isinstance(mod, operators.BuiltParametrizedOperator):
                phi = mod.angle(**kwargs)
                phi = phi.to(device=st.device, dtype=torch.float32)
                if phi.dim() == 1 and phi.numel() == 1 and not mod.named:
                    phi = phi.reshape(())
                if phi.dim() == 1 and st.shape[0] == 1 and phi.shape[0] != 1:
                    st = st.expand(phi.shape[0], -1)
                st = plain.RotFn.apply(st, phi, n, n - 1 - self.inverse[mod.qubit], mod.KIND)
            elif isinstance(mod, operators.BuiltU):
                parent_bit = n - 1 - self.inverse[mod.qubit]
                st = plain.UnitaryFn.apply(st, mod._u.to(st.device), n, parent_bit, 0)
            elif isinstance(mod, measurements.MeasurementDummy):
                st = mod(st)
            else:
                raise NotImplementedError(f"{type(mod).__name__} inside a split sub-circuit")
        return st if batched or st.shape[0] != 1 else st.squeeze(0)

    def to_qasm(self):
        """The layers' statements with sub-circuit qubits renamed to parent qubits
        (``src/qandle/qcircuit.py:297-313``): the ``qubit`` field where a statement
        has one, else every ``q[i]`` in its text."""
        reps = [mod.to_qasm() for mod in self.layers if hasattr(mod, "to_qasm")]
        for r in reps:
            if r.qubit not in ("", None):
                r.qubit = self.inverse[r.qubit]
            else:
                r.gate_str = re.sub(r"q\[(\d+)\]", lambda m: f"q[{self.inverse[int(m.group(1))]}]", r.gate_str)
        return reps
