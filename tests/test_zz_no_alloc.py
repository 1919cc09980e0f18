"""qs_measure_prob_noalloc (partials in a static device array, no workspace
argument) against qs_measure_prob with a caller workspace, through the C-ABI:
bit-identical results on two calls in a row, the second on a smaller input
(the array keeps the first call's partials), and the size limit refused."""
import ctypes

import pytest
import torch

from qandle_amd._engine import lib as L

pytestmark = pytest.mark.gpu


def _state(B, n, seed):
    g = torch.Generator().manual_seed(seed)
    psi = torch.randn(B, 2 ** n, dtype=torch.complex64, generator=g)
    return psi / torch.linalg.norm(psi, dim=-1, keepdim=True)


def _with_workspace(lib, psi, B, n, st):
    out = torch.empty(B, n, dtype=torch.float32, device=psi.device)
    wb = lib.qs_measure_prob_work_bytes(B, n)
    work = torch.empty(max(wb // 4, 1), dtype=torch.float32, device=psi.device)
    L.check(lib.qs_measure_prob(L.ptr(out), L.ptr(psi), B, n, L.ptr(work), wb, st), "qs_measure_prob")
    return out


def _noalloc(lib, psi, B, n, st):
    out = torch.empty(B, n, dtype=torch.float32, device=psi.device)
    L.check(lib.qs_measure_prob_noalloc(L.ptr(out), L.ptr(psi), B, n, st), "qs_measure_prob_noalloc")
    return out


def test_noalloc_matches_workspace_entry_point(device):
    lib = L.load()
    st = torch.cuda.current_stream(device).cuda_stream
    big = _state(16, 20, 1).to(device)          # 2^24 amplitudes: 2^12 chunk partials
    small = _state(3, 13, 2).to(device)         # a smaller second call over the same array
    ref_big = _with_workspace(lib, big, 16, 20, st)
    ref_small = _with_workspace(lib, small, 3, 13, st)
    torch.cuda.synchronize(device)
    free0 = torch.cuda.mem_get_info(device)[0]
    got_big = _noalloc(lib, big, 16, 20, st)
    got_small = _noalloc(lib, small, 3, 13, st)
    torch.cuda.synchronize(device)
    free1 = torch.cuda.mem_get_info(device)[0]
    print(f"device free memory before/after the two calls: {free0} / {free1} (outputs: {16 * 20 * 4 + 3 * 13 * 4} B)")
    assert torch.equal(got_big, ref_big)
    assert torch.equal(got_small, ref_small)
    # a third call of the first size again: the array held the small call's partials
    assert torch.equal(_noalloc(lib, big, 16, 20, st), ref_big)


def test_noalloc_refuses_oversized_requests(device):
    lib = L.load()
    st = torch.cuda.current_stream(device).cuda_stream
    out = torch.empty(4, dtype=torch.float32, device=device)
    psi = torch.empty(4, dtype=torch.complex64, device=device)
    # 2^31 amplitudes need 26 MiB of partials, over the 16 MiB array: refused before any launch
    assert lib.qs_measure_prob_work_bytes(1 << 11, 20) > (16 << 20)
    assert lib.qs_measure_prob_noalloc(L.ptr(out), L.ptr(psi), 1 << 11, 20, st) == 3  # QS_EWORKSPACE
    assert lib.qs_measure_prob_noalloc(L.ptr(out), L.ptr(psi), 0, 20, st) == 1       # QS_EINVAL
    assert lib.qs_measure_prob_noalloc(None, L.ptr(psi), 1, 2, st) == 1
