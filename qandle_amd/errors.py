"""Error types of the public API (``src/qandle/errors.py:1-6``)."""


class UnbuiltGateError(ValueError):
    """Raised when an operator that still needs a parameter or a build is used.

    Mirrors ``qandle.errors.UnbuiltGateError`` (``src/qandle/errors.py:1-6``).
    """


class EngineError(RuntimeError):
    """A non-zero status returned by the HIP engine through its C-ABI."""
