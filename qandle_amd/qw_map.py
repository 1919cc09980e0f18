"""Weight remappings used as rotation-angle maps.

qandle imports these from the third-party package qW-Map 0.1.2
(``src/qandle/config.py:1-3``, ``src/qandle/operators.py:168-171``), which is not
part of this repository.  ``none`` is the identity.  ``tanh`` restates the
published qW-Map rule phi = pi * tanh(theta) (arXiv 2212.14807); no file of the
reference pins its value, so it is documented as unpinned and every parity
workload passes ``remapping=None``.
"""
import math

import torch


def none(x: torch.Tensor) -> torch.Tensor:
    return x


def tanh(x: torch.Tensor) -> torch.Tensor:
    return torch.tanh(x) * math.pi
