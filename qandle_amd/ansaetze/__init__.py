"""Ansätze of the drop-in API (``src/qandle/ansaetze/__init__.py``).

Each ansatz builds the same operator modules (same parameter draws, same
``state_dict`` layout) as the reference.  StronglyEntanglingLayer and
StronglyEntanglingLayerBudget are plain gate sequences, so a circuit holding
them lowers into the fused program; TwoLocal and SpecialUnitary keep the
reference's own forward (see their classes).
"""
from .ansatz import Ansatz, UnbuiltAnsatz, BuiltAnsatz  # noqa: F401
from .stronglyentangling import StronglyEntanglingLayer, StronglyEntanglingLayerBuilt  # noqa: F401
from .twolocal import TwoLocal, TwoLocalBuilt  # noqa: F401
from .specialunitary import SpecialUnitary, SpecialUnitaryBuilt  # noqa: F401
from .stronglyentangling_budget import StronglyEntanglingLayerBudget  # noqa: F401

SU = SpecialUnitary  # alias, as in the reference

__all__ = ["StronglyEntanglingLayer", "TwoLocal", "SpecialUnitary", "SU", "StronglyEntanglingLayerBudget"]
