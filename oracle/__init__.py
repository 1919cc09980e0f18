"""CPU oracle for the qandle gate-application path -- TEST INFRASTRUCTURE ONLY.

Nothing in ``qandle_amd`` imports this package.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker (or, for ``cpu_baseline``, as the timed CPU port of
qandle's own torch path), never as the thing measured on the GPU.

Modules
-------
fp64
    complex128 numpy state-vector simulator written from the closed forms of
    qandle's gates (``src/qandle/operators.py:368-395``), its state layout
    (qubit q is amplitude bit n-1-q, ``src/qandle/operators.py:284-286``), its
    AngleEmbedding (``src/qandle/embeddings.py:148-164``) and MeasureProbability
    (``src/qandle/measurements.py:101-123``), plus a dense-unitary route, an
    FP64 adjoint gradient, central finite differences and parameter shift.
qandle_c64
    restatement of qandle's complex64 *torch* semantics: a ``dense`` mode that
    performs the reference's own ``state @ M`` with the cached 2^n x 2^n
    partial matrices, and a ``structured`` mode that applies the same gates as
    2x2 complex64 updates (for sizes where the dense cache cannot exist).

Pinning (how each oracle function is checked independently, see
``tests/test_oracle.py``):

* gate closed forms   -> ``scipy.linalg.expm(-i phi/2 P)`` (textbook route)
* structured apply    -> explicit Kronecker-product unitaries (n <= 8)
* CNOT                -> brute-force permutation matrix, involution
* AngleEmbedding      -> product of single-qubit expm columns; MeasureProbability
                         of the embedding alone equals cos^2(x/2) (RX/RY)
* MeasureProbability  -> definition by explicit index loop; P0 + P1 = ||psi||^2
* FP64 adjoint        -> central finite differences and parameter shift
* qandle_c64 dense    -> fp64 oracle within complex64 rounding, and dense ==
                         structured at n <= 8
* qw_map.tanh remap   -> parity unpinned (third-party qW-Map 0.1.2 is absent);
                         every parity workload uses ``remapping=None``.
"""
