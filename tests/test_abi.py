"""The C-ABI library loads and exports every entry point include/qandle_sv.h declares.

Only host-side calls are made here (version, status strings, workspace sizes
and argument validation that returns before any HIP call).
"""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qandle_sv.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qs_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from qandle_amd._engine import lib as L

    if not os.path.exists(L.LIB_PATH):
        pytest.fail(f"engine library not built at {L.LIB_PATH}")
    return L.load()


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_table_matches_header():
    from qandle_amd._engine import lib as L

    assert sorted(L.SIGNATURES) == declared_functions()


def test_host_side_calls(lib):
    assert lib.qs_abi_version() >= 1
    assert lib.qs_status_string(0) == b"ok"
    assert lib.qs_status_string(1) == b"invalid argument"
    # (B=4, n=10): one chunk of 2^10 per sample, 11 floats each
    assert lib.qs_measure_prob_work_bytes(4, 10) == 4 * 11 * 4
    assert lib.qs_measure_prob_work_bytes(2, 20) == 2 * 2 ** 8 * 13 * 4
    assert lib.qs_pauli_im_work_bytes(1, 13, 3) == 2 * 3 * 4
    assert lib.qs_measure_prob_work_bytes(0, 3) == -1


def test_argument_validation_returns_einval(lib):
    dummy = ctypes.c_void_p(16)
    # bit out of range, bad kind, bad control == target, null pointers
    assert lib.qs_rot(dummy, dummy, 1, 3, 3, 0, dummy, 0, 0, None) == 1
    assert lib.qs_rot(dummy, dummy, 1, 3, 0, 7, dummy, 0, 0, None) == 1
    assert lib.qs_rot(None, dummy, 1, 3, 0, 0, dummy, 0, 0, None) == 1
    assert lib.qs_cnot(dummy, dummy, 1, 3, 1, 1, None) == 1
    assert lib.qs_cnot(dummy, dummy, 0, 3, 0, 1, None) == 1
    bits = (ctypes.c_int32 * 2)(0, 0)
    assert lib.qs_embed_angle(dummy, dummy, 1, 3, bits, 2, 0, 2, None) == 1  # duplicate bit
    # workspace too small
    assert lib.qs_measure_prob(dummy, dummy, 1, 12, dummy, 4, None) == 3
