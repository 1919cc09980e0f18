"""Lane layout of the wave reductions the adjoint kernels use (DPP exchanges,
row rotations and the gfx950 row/half swaps), checked against host sums."""
import ctypes
import struct

import numpy as np
import pytest
import torch

PROBE = r"""
extern "C" __global__ void __launch_bounds__(64) qs_lane_probe(const float *x, float *out) {
  const int l = threadIdx.x;
  const float a = x[l], b = x[64 + l], c = x[128 + l], d = x[192 + l];
  out[l] = swap16_add(a);
  out[64 + l] = swap32_add(a);
  out[128 + l] = wave_sum4(a, b, c, d);
}
"""


@pytest.mark.gpu
def test_swaps_and_wave_sum4_lane_layout(device):
    from qandle_amd._engine import jit
    from qandle_amd._engine import lib as L

    fn = jit.kernels_for([PROBE], ["qs_lane_probe"])[0]
    gen = torch.Generator().manual_seed(5)
    # small integers: every sum is exact in float32, whatever the order
    x = torch.randint(-8, 9, (256,), generator=gen).float()
    xd = x.to(device)
    out = torch.full((192,), float("nan"), device=device)
    args = struct.pack("<QQ", xd.data_ptr(), out.data_ptr())
    buf = ctypes.create_string_buffer(args, len(args))
    stream = torch.cuda.current_stream(device)
    L.check(L.load().qs_jit_launch(fn, 1, 64, buf, len(args), stream.cuda_stream), "qs_jit_launch")
    torch.cuda.synchronize(device)
    o = out.cpu().numpy()
    a = x.numpy()[:64]
    lanes = np.arange(64)
    assert np.array_equal(o[:64], a + a[lanes ^ 16])
    assert np.array_equal(o[64:128], a + a[lanes ^ 32])
    v = x.numpy().reshape(4, 64).sum(1)
    expect = v[2 * (lanes & 1) + ((lanes >> 1) & 1)]
    assert np.array_equal(o[128:], expect)
