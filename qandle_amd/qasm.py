"""OpenQASM 2 / 3 text export (``src/qandle/qasm.py:1-70``).

Every operator's ``to_qasm`` returns one ``QasmRepresentation`` or a list of
them (nested one level at most, as a split circuit's sub-circuits are); the
text format is the reference's, line for line, so that exported circuits read
the same in either package:

- one statement per gate: ``<gate_str>(<value or qasm3 input>) q[<qubit>];``;
- OpenQASM 3 declares ``input float <name>;`` for every named (input-bound)
  rotation and ``output float <name>;`` for every measured qubit, before the
  gates.
"""
from __future__ import annotations

import dataclasses
import typing

__all__ = ["QasmRepresentation", "convert_to_qasm"]


@dataclasses.dataclass
class QasmRepresentation:
    """One gate statement (``src/qandle/qasm.py:5-30``)."""

    gate_str: str
    gate_value: typing.Union[str, float] = ""
    qubit: typing.Union[str, int] = ""
    qasm3_inputs: str = ""
    qasm3_outputs: str = ""

    def __iter__(self):
        """A single statement iterates as a one-element list."""
        yield self

    def to_string(self) -> str:
        parts = [self.gate_str]
        arg = self.qasm3_inputs if self.qasm3_inputs != "" else self.gate_value
        if arg != "":
            parts.append(f"({arg})")
        if self.qubit not in ("", None):
            parts.append(f" q[{self.qubit}]")
        return "".join(parts) + ";"


def _header(version: int, n: int) -> str:
    if version == 2:
        return f'OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[{n}];\ncreg c[{n}];\n'
    return f"OPENQASM 3.0;\nqubit[{n}] q;\nbit[{n}] c;\n"


def convert_to_qasm(circuit, qasm_version: int = 2, include_header: bool = True) -> str:
    """The circuit as OpenQASM text (``src/qandle/qasm.py:33-70``); any version
    other than 2 writes OpenQASM 3."""
    statements = [r for rep in circuit.to_qasm() for r in rep]
    lines = [_header(qasm_version, circuit.num_qubits)] if include_header else []
    if qasm_version == 3:
        for r in statements:
            if r.qasm3_inputs != "":
                lines.append(f"input float {r.qasm3_inputs};\n")
            if r.qasm3_outputs != "":
                lines.append(f"output float {r.qasm3_outputs};\n")
    lines += [r.to_string() + "\n" for r in statements]
    return "".join(lines)
