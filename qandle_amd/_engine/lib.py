"""ctypes binding of libqandle_sv.so (the C-ABI declared in include/qandle_sv.h).

The library is built in-tree by ``qandle_amd/csrc/build.py`` (hipcc for
gfx950).  There is no fallback: if the library is missing, or a tensor is not
on a HIP device, the call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch

from ..errors import EngineError

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "_lib", "libqandle_sv.so"))

_lock = threading.Lock()
_lib = None

P = ctypes.c_void_p
I32 = ctypes.c_int
I64 = ctypes.c_int64

# name -> (restype, argtypes); must match include/qandle_sv.h
SIGNATURES = {
    "qs_abi_version": (I32, []),
    "qs_last_hip_error": (I32, []),
    "qs_status_string": (ctypes.c_char_p, [I32]),
    "qs_half_angle_coeffs": (I32, [P, P, I64, P]),
    "qs_embed_angle": (I32, [P, P, I64, I32, P, I32, I32, I64, P]),
    "qs_rot": (I32, [P, P, I64, I32, I32, I32, P, I64, I32, P]),
    "qs_cnot": (I32, [P, P, I64, I32, I32, I32, P]),
    "qs_measure_prob_work_bytes": (I64, [I64, I32]),
    "qs_measure_prob": (I32, [P, P, I64, I32, P, I64, P]),
    "qs_measure_prob_noalloc": (I32, [P, P, I64, I32, P]),
    "qs_measure_prob_backward": (I32, [P, P, P, I64, I64, I32, P]),
    "qs_pauli_im_work_bytes": (I64, [I64, I32, I32]),
    "qs_pauli_im": (I32, [P, P, P, I64, I32, P, I32, I32, P, I64, P]),
    "qs_pauli_im_controlled": (I32, [P, P, P, I64, I32, P, I32, I32, I64, P, I64, P]),
    "qs_unitary_1q": (I32, [P, P, I64, I32, I32, P, I64, I64, I32, P]),
    "qs_rot_unitaries": (I32, [P, P, I64, I32, P]),
    "qs_mcx": (I32, [P, P, I64, I32, I32, I64, P]),
    "qs_phase_flip": (I32, [P, P, I64, I32, I64, P]),
    "qs_swap": (I32, [P, P, I64, I32, I32, I32, I64, P]),
    "qs_reset": (I32, [P, P, I64, I32, I32, P]),
    "qs_reset_backward": (I32, [P, P, P, I64, I32, I32, P]),
    "qs_amplitude_embed": (I32, [P, P, I64, I32, I32, I64, ctypes.c_float, I32, P]),
    "qs_amplitude_embed_backward": (I32, [P, P, P, I64, I32, I32, I64, ctypes.c_float, I32, P]),
    "qs_abs2": (I32, [P, P, I64, P]),
    "qs_abs2_backward": (I32, [P, P, P, I64, P]),
    "qs_fused_supported": (I32, [I32, I32]),
    "qs_layout_supported": (I32, [I32, I32]),
    "qs_product_factors": (I32, [P, P, I64, I32, P, I32, I32, I64, P]),
    "qs_factor_apply": (I32, [P, P, P, I64, P, P, I64, I32, P]),
    "qs_fused_pass": (I32, [P, P, I64, I32, I32, I32, P, P, I32, P, P, P, P, I32, P]),
    "qs_jit_compile": (I32, [ctypes.c_char_p, ctypes.c_char_p, P, I32, P]),
    "qs_jit_code_size": (I64, [P]),
    "qs_jit_code_copy": (I32, [P, P, I64]),
    "qs_jit_log_size": (I64, [P]),
    "qs_jit_log_copy": (I32, [P, P, I64]),
    "qs_jit_release_code": (None, [P]),
    "qs_jit_load": (I32, [P, I64, P]),
    "qs_jit_function": (I32, [P, ctypes.c_char_p, P]),
    "qs_jit_resident_blocks": (I64, [P, I32]),
    "qs_jit_launch": (I32, [P, ctypes.c_uint, ctypes.c_uint, P, I64, P]),
    "qs_jit_unload": (I32, [P]),
    "qs_compose_tables": (I32, [P, P, P, P, I32, P, P]),
    "qs_fused_measure_work_bytes": (I64, [I64, I32, I32, I32]),
    "qs_fused_measure_finish": (I32, [P, P, I64, I32, I32, I32, P, P]),
    "qs_readout_work_bytes": (I64, [I64, I32, I32, I32]),
    "qs_lane_coeffs": (I32, [P, P, P, I64, I32, P]),
    "qs_readout_finish": (I32, [P, P, I64, I32, I32, I32, P]),
    "qs_adjoint_work_bytes": (I64, [I64, I32, I32, I32, I32]),
    "qs_adjoint_pass": (I32, [P, P, I64, I32, I32, I32, P, P, I32, P, P, P, P, I32, P, P]),
    "qs_adjoint_finish": (I32, [P, P, P, I32, I64, I32, I32, I32, P]),
    "qs_factor_entries": (I32, [P, P, P, I64, P]),
    "qs_pass_desc_bytes": (I64, []),
    "qs_pass_desc": (I32, [P, P, P, P, I64, I32, I32, I32, P, P, I32, P, P, P, P, I32]),
    "qs_items_pass": (I32, [P, P, I32, I32, I32, I32, P]),
    "qs_measure_desc_bytes": (I64, []),
    "qs_measure_desc": (I32, [P, P, P, I64, I32, I32, I32, P]),
    "qs_items_measure_finish": (I32, [P, I32, I32, P]),
    "qs_adjoint_desc_bytes": (I64, []),
    "qs_adjoint_desc": (I32, [P, P, P, P, I64, I32, I32, I32, P, P, I32, P, P, P, P, I32, P]),
    "qs_items_adjoint": (I32, [P, P, I32, I32, I32, I32, P]),
    "qs_adjoint_finish_desc_bytes": (I64, []),
    "qs_adjoint_finish_desc": (I32, [P, P, P, P, I32, I64, I32, I32, I32]),
    "qs_items_adjoint_finish": (I32, [P, I32, I32, P]),
}


def load():
    """Load (once) and return the engine library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise EngineError(
                f"HIP engine library not found at {LIB_PATH}; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)"
            )
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(status: int, what: str) -> None:
    if status != 0:
        lib = load()
        msg = lib.qs_status_string(status).decode()
        extra = f" (hip error {lib.qs_last_hip_error()})" if status == 2 else ""
        raise EngineError(f"{what}: {msg}{extra}")


def stream_of(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def require_device(t: torch.Tensor, what: str = "tensor") -> None:
    if t.device.type != "cuda":
        raise EngineError(
            f"qandle_amd runs on the MI355X HIP engine only; {what} is on {t.device}. "
            "Move the circuit and its inputs to a HIP device (module.to('cuda'))."
        )


def ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


_copy_streams: dict = {}


def upload(values, dtype, device) -> torch.Tensor:
    """Host values (a list or numpy array) as a tensor on ``device``, resident
    when this returns.

    A copy from pageable memory on the current stream returns only once every
    kernel queued before it on that stream has run: a serving loop that lowers
    a new circuit while its stream is busy would stall on it.  Here the bytes
    go through pinned host memory on a stream of their own (default priority,
    nothing else queued on it), and the host waits for that copy alone.  The
    tensor is allocated on the copy stream and recorded on the current one,
    so its memory is not reused before the current stream's work at its
    release has run."""
    t = torch.as_tensor(values, dtype=dtype)
    dev = torch.device(device)
    if dev.type != "cuda":
        return t
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    with _lock:
        st = _copy_streams.get(dev.index)
        if st is None:
            st = _copy_streams[dev.index] = torch.cuda.Stream(dev)
    src = t.pin_memory()
    cur = torch.cuda.current_stream(dev)
    with torch.cuda.stream(st):
        out = torch.empty(t.shape, dtype=dtype, device=dev)
        out.copy_(src, non_blocking=True)
    out.record_stream(cur)
    st.synchronize()
    return out


def i32_array(values):
    values = [int(v) for v in values]
    return (ctypes.c_int32 * max(1, len(values)))(*values)
