"""qandle_amd: qandle's circuit API on a hand-written MI355X (gfx950) HIP engine.

Public names follow ``src/qandle/__init__.py:3-14`` for the gate-application
path: ``Circuit``, ``RX``/``RY``/``RZ``, ``CNOT``, ``AngleEmbedding``,
``MeasureProbability``, ``UnbuiltGateError``, ``config``, ``split``, and the
host-side ``draw`` and ``convert_to_qasm`` / ``QasmRepresentation``.
``import qandle_amd as qandle`` is the intended drop-in spelling.
"""
from . import config, errors, qw_map, utils  # noqa: F401
from .errors import UnbuiltGateError, EngineError  # noqa: F401
from .operators import *  # noqa: F401,F403
from .operators import BuiltRX, BuiltRY, BuiltRZ, BuiltCNOT  # noqa: F401
from .embeddings import *  # noqa: F401,F403
from .measurements import *  # noqa: F401,F403
from .qcircuit import *  # noqa: F401,F403
from .splitter import split, SubcircuitContainer  # noqa: F401
from .ansaetze import *  # noqa: F401,F403
from .convolution import QConv  # noqa: F401
from .drawer import *  # noqa: F401,F403
from .qasm import *  # noqa: F401,F403
from . import operators, embeddings, measurements, qcircuit, splitter, ansaetze, convolution  # noqa: F401
from . import drawer, qasm  # noqa: F401

__version__ = "0.1.0"
