"""Per-pass specialised forward kernels, compiled at run time.

The interpreted pass kernel (``qs_fused_pass``) reads its program words while
it runs: every segment head, table index and CNOT costs a dependent scalar
load and a branch, and the few waves a 2^11-amplitude tile allows cannot hide
that latency.  Here the same encoded words of a pass are turned into the HIP
source of one kernel with everything constant:

* the tile layout (local and non-local bit positions, the LDS swizzle strides
  of every register), the sub-pass register/thread maps;
* fused unitaries read at constant table offsets (scalar loads, issued at
  once) and applied on the matrix cores (``v_mfma_f32_4x4x1f32``);
* CNOTs between register bits become a renaming of registers at code
  generation time (no instruction at all); CNOTs with a control outside the
  registers become per-thread selects;
* the RZ phases, the product-state generator and the MeasureProbability
  partials with their positions fixed.

The kernels of one circuit's passes form one program: hipRTC compiles one
program at a time, and the compiler start-up is paid once per program.

The kernel keeps the interpreted kernel's tile pipeline (the next tile's HBM
loads in flight during the current tile's arithmetic), its LDS layout, its
table formats (``qs_compose_tables``) and its measurement partial layout
(``qs_fused_measure_finish``), so both paths produce the same results on the
same inputs.  Sources are compiled once per process with hipRTC through the
C-ABI (``qs_jit_*``) and the code objects are cached on disk by source hash.
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import struct
import threading

from . import lib as L
from . import planner as PL

PRELUDE = r"""
typedef long long i64;
typedef float qs_f4 __attribute__((ext_vector_type(4)));
struct JArgs {
  float *state; const float *src; const int *utab; const int *cs; const float *efac; float *mpart;
  const float *ltab;
  i64 B; i64 ntiles; int store; int pad;
};
__device__ __forceinline__ int swz(int l) { int f = (l >> 5) ^ (l >> 10) ^ (l >> 15); return l ^ (f & 31); }
// the swizzle with bit 0 left alone: an amplitude pair keeps one 16-byte slot in order
__device__ __forceinline__ int swz_e(int l) { int f = (l >> 5) ^ (l >> 10) ^ (l >> 15); return l ^ (f & 30); }
// a tile element by byte offset: with the offset kept as (index << 3) the swizzled address of
// register j is one XOR with a constant, (lb << 3) ^ (stride_j << 3)
__device__ __forceinline__ float2 &lds_at(float2 *base, int byte_off) {
  return *reinterpret_cast<float2 *>(reinterpret_cast<char *>(base) + byte_off);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// the composed tables are read-only while a pass runs: reading them through the
// constant address space lets uniform reads use the scalar (SMEM) path
typedef const __attribute__((address_space(4))) int cint;
__device__ __forceinline__ cint *as_const(const int *p) { return (cint *)(p); }
// an opaque copy of a table pointer: the loads of one unitary's block cannot
// be hoisted above it, so only that block's matrices occupy scalar registers
// (hoisting every table read of a pass to the kernel entry runs out of SGPRs)
__device__ __forceinline__ const int *pin(const int *p) { asm volatile("" : "+s"(p)); return p; }
__device__ __forceinline__ float *pinf(float *p) { asm volatile("" : "+s"(p)); return p; }
// an opaque copy of the thread index: addresses derived from it inside the
// tile loop are recomputed where they are used instead of being hoisted out
// of the loop and kept live in vector registers for the whole kernel
__device__ __forceinline__ int pinv(int x) { asm volatile("" : "+v"(x)); return x; }
// lane coefficients of a fused unitary (8 floats as int bits at u) for
// v_mfma_f32_4x4x1f32: A_k = U[lane & 3][k] of the real 4x4 form
__device__ __forceinline__ void coeffs(const int *up, float (&m)[4]) {
  cint *u = as_const(up);
  const float a = __int_as_float(u[0]), b = __int_as_float(u[1]), c = __int_as_float(u[2]),
              d = __int_as_float(u[3]), e = __int_as_float(u[4]), f = __int_as_float(u[5]),
              g = __int_as_float(u[6]), h = __int_as_float(u[7]);
  const int row = threadIdx.x & 3;
  const bool r1 = row & 1, r2 = row & 2;
  m[0] = r2 ? (r1 ? f : e) : (r1 ? b : a);
  m[1] = r2 ? (r1 ? e : -f) : (r1 ? a : -b);
  m[2] = r2 ? (r1 ? h : g) : (r1 ? d : c);
  m[3] = r2 ? (r1 ? g : -h) : (r1 ? c : -d);
}
__device__ __forceinline__ void mfma_pair(float2 &x, float2 &y, const float (&m)[4]) {
  qs_f4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_4x4x1f32(m[0], x.x, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_4x4x1f32(m[1], x.y, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_4x4x1f32(m[2], y.x, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_4x4x1f32(m[3], y.y, acc, 0, 0, 0);
  x = make_float2(acc[0], acc[1]);
  y = make_float2(acc[2], acc[3]);
}
// two independent pairs with their MFMA chains interleaved: each dependent
// accumulation then finds the other chain's MFMA between it and its producer
// instead of wait states
__device__ __forceinline__ void mfma_pair2(float2 &x0, float2 &y0, float2 &x1, float2 &y1, const float (&m)[4]) {
  qs_f4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f};
  a0 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[0], x0.x, a0, 0, 0, 0);
  a1 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[0], x1.x, a1, 0, 0, 0);
  a0 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[1], x0.y, a0, 0, 0, 0);
  a1 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[1], x1.y, a1, 0, 0, 0);
  a0 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[2], y0.x, a0, 0, 0, 0);
  a1 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[2], y1.x, a1, 0, 0, 0);
  a0 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[3], y0.y, a0, 0, 0, 0);
  a1 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[3], y1.y, a1, 0, 0, 0);
  x0 = make_float2(a0[0], a0[1]);
  y0 = make_float2(a0[2], a0[3]);
  x1 = make_float2(a1[0], a1[1]);
  y1 = make_float2(a1[2], a1[3]);
}
// four independent pairs, their MFMA chains interleaved: three other MFMAs
// separate every dependent accumulation from its producer
__device__ __forceinline__ void mfma_pair4(float2 &x0, float2 &y0, float2 &x1, float2 &y1, float2 &x2, float2 &y2,
                                           float2 &x3, float2 &y3, const float (&m)[4]) {
  qs_f4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f}, a2 = {0.f, 0.f, 0.f, 0.f}, a3 = {0.f, 0.f, 0.f, 0.f};
  a0 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[0], x0.x, a0, 0, 0, 0);
  a1 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[0], x1.x, a1, 0, 0, 0);
  a2 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[0], x2.x, a2, 0, 0, 0);
  a3 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[0], x3.x, a3, 0, 0, 0);
  a0 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[1], x0.y, a0, 0, 0, 0);
  a1 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[1], x1.y, a1, 0, 0, 0);
  a2 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[1], x2.y, a2, 0, 0, 0);
  a3 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[1], x3.y, a3, 0, 0, 0);
  a0 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[2], y0.x, a0, 0, 0, 0);
  a1 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[2], y1.x, a1, 0, 0, 0);
  a2 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[2], y2.x, a2, 0, 0, 0);
  a3 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[2], y3.x, a3, 0, 0, 0);
  a0 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[3], y0.y, a0, 0, 0, 0);
  a1 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[3], y1.y, a1, 0, 0, 0);
  a2 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[3], y2.y, a2, 0, 0, 0);
  a3 = __builtin_amdgcn_mfma_f32_4x4x1f32(m[3], y3.y, a3, 0, 0, 0);
  x0 = make_float2(a0[0], a0[1]); y0 = make_float2(a0[2], a0[3]);
  x1 = make_float2(a1[0], a1[1]); y1 = make_float2(a1[2], a1[3]);
  x2 = make_float2(a2[0], a2[1]); y2 = make_float2(a2[2], a2[3]);
  x3 = make_float2(a3[0], a3[1]); y3 = make_float2(a3[2], a3[3]);
}
// coefficients of U^dagger (the transposed real 4x4 form): A_k = U[k][lane & 3]
__device__ __forceinline__ void coeffs_dag(const int *up, float (&m)[4]) {
  cint *u = as_const(up);
  const float a = __int_as_float(u[0]), b = __int_as_float(u[1]), c = __int_as_float(u[2]),
              d = __int_as_float(u[3]), e = __int_as_float(u[4]), f = __int_as_float(u[5]),
              g = __int_as_float(u[6]), h = __int_as_float(u[7]);
  const int row = threadIdx.x & 3;
  const bool r1 = row & 1, r2 = row & 2;
  m[0] = r2 ? (r1 ? -d : c) : (r1 ? -b : a);
  m[1] = r2 ? (r1 ? c : d) : (r1 ? a : b);
  m[2] = r2 ? (r1 ? -h : g) : (r1 ? -f : e);
  m[3] = r2 ? (r1 ? g : h) : (r1 ? e : f);
}
// U^dagger on the vector ALU
__device__ __forceinline__ void valu_pair_dag(float2 &x, float2 &y, const int *u) {
  const float a = __int_as_float(u[0]), b = -__int_as_float(u[1]), c = __int_as_float(u[4]),
              d = -__int_as_float(u[5]), e = __int_as_float(u[2]), f = -__int_as_float(u[3]),
              g = __int_as_float(u[6]), h = -__int_as_float(u[7]);
  const float2 p = x, q = y;
  x = make_float2(a * p.x - b * p.y + c * q.x - d * q.y, a * p.y + b * p.x + c * q.y + d * q.x);
  y = make_float2(e * p.x - f * p.y + g * q.x - h * q.y, e * p.y + f * p.x + g * q.y + h * q.x);
}
// Im Tr(Q M) for a gradient generator Q (8 floats as int bits) and a cross density M
__device__ __forceinline__ float im_trace(const int *qp, const float (&m)[8]) {
  cint *q = as_const(qp);
  return (__int_as_float(q[0]) * m[1] + __int_as_float(q[1]) * m[0]) +
         (__int_as_float(q[2]) * m[5] + __int_as_float(q[3]) * m[4]) +
         (__int_as_float(q[4]) * m[3] + __int_as_float(q[5]) * m[2]) +
         (__int_as_float(q[6]) * m[7] + __int_as_float(q[7]) * m[6]);
}
// Im Tr(Q M) for a Hermitian generator Q (V P V^dagger): with Q00, Q11 real and
// Q01 = conj(Q10) it needs only Im M00, Im M11, Im M01 + Im M10 and Re M01 - Re M10
__device__ __forceinline__ float herm_trace(const int *qp, const float (&h)[4]) {
  cint *q = as_const(qp);
  return (__int_as_float(q[0]) * h[0] + __int_as_float(q[6]) * h[1]) +
         (__int_as_float(q[4]) * h[2] + __int_as_float(q[5]) * h[3]);
}
// sums over lane groups with DPP (no LDS traffic): quad_perm xor 1 and xor 2,
// then row_ror 4 and 8 (every lane holds its 16-lane row sum)
template <int CTRL>
__device__ __forceinline__ float dpp_add(float s) {
  return s + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(s), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float quad_sum(float s) { return dpp_add<0x4E>(dpp_add<0xB1>(s)); }
__device__ __forceinline__ float row_sum(float s) { return dpp_add<0x128>(dpp_add<0x124>(quad_sum(s))); }
__device__ __forceinline__ float lane_f(float s, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(s), l));
}
// 64-lane sum (uniform)
__device__ __forceinline__ float wave_sum(float s) {
  const float r = row_sum(s);
  return (lane_f(r, 0) + lane_f(r, 16)) + (lane_f(r, 32) + lane_f(r, 48));
}
// 32-lane half sums (each lane gets its half's sum)
__device__ __forceinline__ float half_sum(float s) {
  const float r = row_sum(s);
  const float lo = lane_f(r, 0) + lane_f(r, 16), hi = lane_f(r, 32) + lane_f(r, 48);
  return (threadIdx.x & 32) ? hi : lo;
}
// Lane l + 16 (row swap, permlane16_swap) and lane l + 32 (half swap, permlane32_swap) folded in:
// with the same value as both operands, the two results of a swap are the value's two halves, so
// their sum is s[l] + s[l ^ 16] (s[l] + s[l ^ 32]) in every lane.
__device__ __forceinline__ float swap16_add(float s) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(s), __float_as_uint(s), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float swap32_add(float s) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(s), __float_as_uint(s), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
// Four values summed over a 64-lane wave together: two exchange rounds (lane l ^ 1, then l ^ 2) leave
// each lane with one value's quad sum, value index 2 * (l & 1) + ((l >> 1) & 1); row_ror 4 and 8 add
// the row's lanes of the same index, the two swaps the four rows.  Every lane ends with the wave sum
// of its value index (lanes 0, 2, 1, 3 hold values 0, 1, 2, 3).
__device__ __forceinline__ float wave_sum4(float t0, float t1, float t2, float t3) {
  const int l = threadIdx.x;
  const bool b0 = l & 1, b1 = l & 2;
  const float k0 = b0 ? t2 : t0, k1 = b0 ? t3 : t1, s0 = b0 ? t0 : t2, s1 = b0 ? t1 : t3;
  const float u0 = k0 + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(s0), 0xB1, 0xF, 0xF, false));
  const float u1 = k1 + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(s1), 0xB1, 0xF, 0xF, false));
  const float k = b1 ? u1 : u0, s = b1 ? u0 : u1;
  float w = k + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(s), 0x4E, 0xF, 0xF, false));
  w = dpp_add<0x128>(dpp_add<0x124>(w));
  return swap32_add(swap16_add(w));
}
struct JAdj {
  float *psi; float *lam; const int *utab; const int *qtab; const int *cs; float *gpart; const float *seed_w;
  const float *ltab;
  const float *psi_src;  // where the step reads psi (psi itself, or a checkpoint): it writes psi
  i64 B; i64 NS; i64 ntiles;
};
// vector-ALU form with the unitary read through the scalar path (uniform tables)
__device__ __forceinline__ void valu_pair_s(float2 &x, float2 &y, const int *up) {
  cint *u = as_const(up);
  const float a = __int_as_float(u[0]), b = __int_as_float(u[1]), c = __int_as_float(u[2]),
              d = __int_as_float(u[3]), e = __int_as_float(u[4]), f = __int_as_float(u[5]),
              g = __int_as_float(u[6]), h = __int_as_float(u[7]);
  const float2 p = x, q = y;
  x = make_float2(a * p.x - b * p.y + c * q.x - d * q.y, a * p.y + b * p.x + c * q.y + d * q.x);
  y = make_float2(e * p.x - f * p.y + g * q.x - h * q.y, e * p.y + f * p.x + g * q.y + h * q.x);
}
// vector-ALU form (per-sample unitaries that may differ inside a 4-lane block)
__device__ __forceinline__ void valu_pair(float2 &x, float2 &y, const int *u) {
  const float a = __int_as_float(u[0]), b = __int_as_float(u[1]), c = __int_as_float(u[2]),
              d = __int_as_float(u[3]), e = __int_as_float(u[4]), f = __int_as_float(u[5]),
              g = __int_as_float(u[6]), h = __int_as_float(u[7]);
  const float2 p = x, q = y;
  x = make_float2(a * p.x - b * p.y + c * q.x - d * q.y, a * p.y + b * p.x + c * q.y + d * q.x);
  y = make_float2(e * p.x - f * p.y + g * q.x - h * q.y, e * p.y + f * p.x + g * q.y + h * q.x);
}
"""

ARGS_FMT = "<QQQQQQQqqii"  # struct JArgs
ADJ_FMT = "<QQQQQQQQQqqq"  # struct JAdj
OPTIONS = ["--offload-arch=gfx950", "-O3", "-std=c++17",
           "-mllvm", "-simplifycfg-sink-common=false",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _swz(l: int, mask: int = 31) -> int:
    return PL.swz(l, mask)


def lane_entries(words, R: int, TB: int, multi: bool):
    """utab indices of the uniform (matrix-core) unitaries of a pass, distinct,
    in program order: the pass's block of the lane-coefficient table."""
    out = []
    for sub in PL.decode_pass(words, R, TB):
        for bits, _c in sub.segments:
            for _r, ui, ps in bits:
                if not ps and ui not in out:
                    out.append(int(ui))
    return out


def pass_source(name: str, words, n: int, R: int, TB: int, lpos, gen: bool, measure: bool,
                prefetch: bool = False, waves: int = None, readout=None, valu_pairs: int = 0,
                lanes=False, split_pairs: bool = False, lane_off: int = 0, interleave: bool = False,
                wave_sync: bool = False, stagger: int = 0, stagger_cu: int = 256, direct_store: bool = False,
                direct_load: bool = False) -> str:
    """HIP source of one specialised forward pass kernel (extern "C" ``name``).

    ``prefetch`` keeps the next tile's HBM loads in flight in registers while
    the current tile computes; ``waves`` caps the registers at that many
    waves per SIMD (None: the compiler's choice).  ``readout``: per bit
    position p, the mask whose parity is the measured bit (the measuring pass
    then writes n + 1 readout partials per slot, qs_readout_finish); None
    measures the bits themselves (K + 1 partials per slot by local bit,
    qs_fused_measure_finish).  ``valu_pairs``: of the amplitude pairs a
    thread updates per uniform fused unitary, that many run on the vector ALU
    and the rest on the matrix cores (the two units then share the work).
    ``lanes``: the matrix-core operands come from the lane-coefficient table
    (``qs_lane_coeffs``) instead of scalar reads and per-lane selects: True
    loads each unitary's operands from it at the start of its sub-pass,
    "lds" copies the pass's block (entries ``lane_off`` onward, in
    ``lane_entries`` order) into LDS with the tile and reads them from there.
    ``split_pairs``: tile loads and stores move each amplitude pair as two
    8-byte LDS accesses (the swizzle may swap a pair's halves) instead of one
    16-byte access with selects.  ``interleave``: the matrix-core updates of a
    lane-table unitary run two (True or 2) or four (4) amplitude pairs at a
    time with their MFMA chains interleaved (mfma_pair2, mfma_pair4), which
    fills the dependent-accumulation wait states; same arithmetic per pair,
    so results are bit for bit equal.  ``stagger``: the second workgroup of
    each compute unit (workgroups ``stagger_cu`` to 2 ``stagger_cu`` - 1)
    sleeps that many s_sleep 127 periods before starting (measured: no effect).
    ``wave_sync``: between two sub-passes that give the thread-index bits
    above the lane bits (the wave index) the same tile bits, each wave works
    on its own part of the tile and a wave-level fence replaces the block
    barrier (wave_local).  ``direct_store``: the last sub-pass stores its
    registers straight to the state (as pass_source_tail does) instead of
    through the tile; ``direct_load``: the first sub-pass loads its registers
    straight from the state (no staging through the tile); both for tiles
    within one sample, and direct_load for one tile per workgroup (no
    prefetch)."""
    K = R + TB
    T, NR = 1 << TB, 1 << R
    multi = n < K
    direct_store = direct_store and not multi
    direct_load = direct_load and not multi and not gen and not prefetch
    subs = PL.decode_pass(words, R, TB)
    lpos = [int(x) for x in lpos]
    npos = [] if multi else [p for p in range(n) if p not in lpos]
    nn = len(npos)
    o = []
    e = o.append
    wattr = f" __attribute__((amdgpu_waves_per_eu({waves})))" if waves else ""
    e(f'extern "C" __global__ void __launch_bounds__({T}){wattr}')
    e(f"{name}(JArgs A) {{")
    e(f"  __shared__ __align__(16) float2 tile[{1 << K}];")
    if measure and TB > 6:
        e(f"  __shared__ float red[{(T // 64) * ((K + 1) if readout is None else (n + 1))}];")
    lane_k = None
    if lanes == "lds":
        ents = lane_entries(words, R, TB, multi)
        lane_k = {ui: k for k, ui in enumerate(ents)}
        if ents:
            e(f"  __shared__ __align__(16) float4 ctab[{4 * len(ents)}];")
    if stagger:
        # the second workgroup of each compute unit starts later, so that the two alternate between
        # streaming their tiles and computing on them instead of doing both in step
        e(f"  if (blockIdx.x >= {stagger_cu} && blockIdx.x < {2 * stagger_cu}) "
          f"for (int i = 0; i < {stagger}; ++i) __builtin_amdgcn_s_sleep(127);")
    e("  const int tid = threadIdx.x;")
    e(f"  const i64 total = A.B << {n};")
    e("  i64 g_tid = 0;")
    for i in range(TB):
        e(f"  g_tid |= (i64)((tid >> {i}) & 1) << {lpos[i + 1]};")
    goff = [sum(1 << lpos[TB + 1 + i] for i in range(R - 1) if (it >> i) & 1) for it in range(NR // 2)]
    e(f"  float4 pre[{max(NR // 2, 1)}];")
    e("  auto tile_base = [&](i64 t) -> i64 {")
    if multi:
        e(f"    return t << {K};")
    else:
        e(f"    i64 b = (t >> {nn}) << {n};")
        for i, p in enumerate(npos):
            e(f"    b |= ((t >> {i}) & 1) << {p};")
        e("    return b;")
    e("  };")
    pf = prefetch and not gen
    if not gen:
        e("  auto fetch = [&](i64 base) {")
        for it in range(NR // 2):
            e(f"    {{ const i64 gi = base | g_tid | {goff[it]}LL;")
            if multi:  # the last tile of a multi-sample pass can run past the batch
                e(f"      pre[{it}] = gi < total ? reinterpret_cast<const float4 *>(A.src)[gi >> 1]"
                  " : make_float4(0.f, 0.f, 0.f, 0.f); }")
            else:
                e(f"      pre[{it}] = reinterpret_cast<const float4 *>(A.src)[gi >> 1]; }}")
        e("  };")
    e("  i64 t = blockIdx.x;")
    if pf:
        e("  if (t < A.ntiles) fetch(tile_base(t));")
    e("  for (; t < A.ntiles; t += gridDim.x) {")
    e("    const i64 base = tile_base(t);")
    if not gen and not pf and not direct_load:
        e("    fetch(base);")
    e("    const int ts = pinv(tid);")
    if lane_k:
        nl = 4 * len(lane_k)
        e(f"    for (int c = ts; c < {nl}; c += {T}) ctab[c] = reinterpret_cast<const float4 *>(A.ltab)[{4 * lane_off} + c];")
        if gen or direct_load:
            e("    __syncthreads();")
    if not gen and not direct_load:
        for it in range(NR // 2):
            if split_pairs:
                e(f"    {{ const int sl = swz(2 * ({it * T} + ts)); const float4 val = pre[{it}];")
                e("      tile[sl] = make_float2(val.x, val.y); tile[sl ^ 1] = make_float2(val.z, val.w); }")
            else:
                e(f"    {{ const int sl = swz(2 * ({it * T} + ts)); const float4 val = pre[{it}]; const bool sw = sl & 1;")
                e("      *reinterpret_cast<float4 *>(&tile[sl & ~1]) = make_float4(sw ? val.z : val.x, sw ? val.w : val.y,"
                  " sw ? val.x : val.z, sw ? val.y : val.w); }")
        e("    __syncthreads();")
        if pf:
            e("    { const i64 tn = t + gridDim.x; if (tn < A.ntiles) fetch(tile_base(tn)); }")
    e(f"    const i64 block_sample = base >> {n};")
    e(f"    float2 v[{NR}];")
    for s, sub in enumerate(subs):
        _sub_pass(e, sub, s, len(subs), n, R, TB, multi, gen and s == 0, measure and s == len(subs) - 1, readout,
                  valu_pairs, lanes=lane_k if lanes == "lds" else lanes, interleave=interleave,
                  wave_sync=wave_sync and s + 1 < len(subs) and wave_local(sub, subs[s + 1], TB),
                  tail=[] if direct_store and s == len(subs) - 1 else None, direct_load=direct_load and s == 0)
    if direct_store:
        e("    __syncthreads();")
        e("  }")
        e("}")
        return "\n".join(o)
    e("    if (A.store) {")
    e("      const int tz = pinv(tid);")
    for it in range(NR // 2):
        e(f"      {{ const i64 gi = base | g_tid | {goff[it]}LL; const int sl = swz(2 * ({it * T} + tz));")
        if split_pairs:
            e("        const float2 lo = tile[sl], hi = tile[sl ^ 1]; const float4 val = make_float4(lo.x, lo.y, hi.x, hi.y);")
        else:
            e("        const float4 raw = *reinterpret_cast<const float4 *>(&tile[sl & ~1]); const bool sw = sl & 1;")
            e("        const float4 val = make_float4(sw ? raw.z : raw.x, sw ? raw.w : raw.y, sw ? raw.x : raw.z,"
              " sw ? raw.y : raw.w);")
        if multi:
            e("        if (gi < total) reinterpret_cast<float4 *>(A.state)[gi >> 1] = val; }")
        else:
            e("        reinterpret_cast<float4 *>(A.state)[gi >> 1] = val; }")
    e("    }")
    e("    __syncthreads();")
    e("  }")
    e("}")
    return "\n".join(o)


def async_load_offsets(lpos, R: int, TB: int):
    """Amplitude offsets of the direct tile loads of pass_source_async.

    16-byte LDS slot s of a tile buffer holds tile elements swz_e(2s) and
    swz_e(2s) + 1 (the pair swz_e keeps in order); load ii of thread tid fills
    slot s = ii * 2^TB + tid (lane j of a wave's 64-lane load lands in slot j of
    its 1 KiB block).  Its global amplitude offset is scatter(swz_e(2s)), where
    scatter puts tile element bit k at amplitude bit lpos[k].  Both maps are
    linear over XOR, so the offset is the XOR of per-thread-bit constants
    ``d_bit[b]`` (bit b of tid set) and a per-load constant ``d_inst[ii]``."""
    K = R + TB
    T = 1 << TB

    def scatter(l):
        return sum(1 << lpos[k] for k in range(K) if (l >> k) & 1)

    d_bit = [scatter(_swz(2 << b, 30)) for b in range(TB)]
    d_inst = [scatter(_swz(2 * ii * T, 30)) for ii in range(1 << (R - 1))]
    return d_bit, d_inst


def pass_source_async(name: str, words, n: int, R: int, TB: int, lpos, measure: bool, waves: int = 2,
                      readout=None, lanes="lds", lane_off: int = 0, interleave: int = 4,
                      wave_sync: bool = True) -> str:
    """HIP source of one specialised forward pass kernel whose HBM reads overlap
    its arithmetic: one persistent workgroup per compute unit holds two tile
    buffers in LDS (2 x 2^(R+TB) amplitudes).  While the sub-passes work on the
    tile in one buffer, the next tile streams from HBM straight into the other
    with gfx950's direct global->LDS loads (``global_load_lds_dwordx4``: no
    registers, no LDS write instructions); the finished tile is stored from LDS
    through registers as in pass_source.  The two buffers are separate LDS
    arrays and the tile loop is unrolled by two, so that the compiler can tell
    the buffer being loaded from the one being read and does not wait for the
    loads before every LDS read.  The tile layout keeps amplitude pairs in order
    in one 16-byte slot (swizzle ``swz_e``), so that lane j of a 64-lane load
    lands in slot j of its 1 KiB block.  Tiles within one sample only
    (n >= R + TB), and not the first pass of a program (it generates its tile
    rather than loading it).  Same sub-pass code, tables and partial layouts as
    pass_source: results are the same bit for bit."""
    K = R + TB
    T, NR = 1 << TB, 1 << R
    if n < K:
        raise ValueError("asynchronous tile loads need tiles within one sample")
    subs = PL.decode_pass(words, R, TB)
    lpos = [int(x) for x in lpos]
    npos = [p for p in range(n) if p not in lpos]
    nn = len(npos)

    o = []
    e = o.append
    wattr = f" __attribute__((amdgpu_waves_per_eu({waves})))" if waves else ""
    e(f'extern "C" __global__ void __launch_bounds__({T}){wattr}')
    e(f"{name}(JArgs A) {{")
    e(f"  __shared__ __align__(16) float2 buf0[{1 << K}];")
    e(f"  __shared__ __align__(16) float2 buf1[{1 << K}];")
    if measure and TB > 6:
        e(f"  __shared__ float red[{(T // 64) * ((K + 1) if readout is None else (n + 1))}];")
    lane_k = None
    if lanes == "lds":
        ents = lane_entries(words, R, TB, False)
        lane_k = {ui: k for k, ui in enumerate(ents)}
        if ents:
            e(f"  __shared__ __align__(16) float4 ctab[{4 * len(ents)}];")
    e("  const int tid = threadIdx.x;")
    e("  i64 g_tid = 0;")
    for i in range(TB):
        e(f"  g_tid |= (i64)((tid >> {i}) & 1) << {lpos[i + 1]};")
    goff = [sum(1 << lpos[TB + 1 + i] for i in range(R - 1) if (it >> i) & 1) for it in range(NR // 2)]
    d_bit, d_inst = async_load_offsets(lpos, R, TB)
    e("  i64 d_tid = 0;")
    for b in range(TB):
        e(f"  d_tid ^= ((tid >> {b}) & 1) ? (i64){d_bit[b]}LL : 0LL;")
    e("  auto tile_base = [&](i64 t) -> i64 {")
    e(f"    i64 b = (t >> {nn}) << {n};")
    for i, p in enumerate(npos):
        e(f"    b |= ((t >> {i}) & 1) << {p};")
    e("    return b;")
    e("  };")
    e("  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);")
    for bsel in (0, 1):
        e(f"  auto issue{bsel} = [&](i64 tt) {{")
        e("    const i64 b = tile_base(tt);")
        e("    const float4 *src = reinterpret_cast<const float4 *>(A.src);")
        for ii in range(NR // 2):
            e(f"    __builtin_amdgcn_global_load_lds(src + ((b | (d_tid ^ {d_inst[ii]}LL)) >> 1),")
            e(f"        (__attribute__((address_space(3))) void *)(buf{bsel} + 2 * ({ii * T} + 64 * wave)), 16, 0, 0);")
        e("  };")
    if lane_k:
        nl = 4 * len(lane_k)
        e(f"  for (int c = tid; c < {nl}; c += {T}) ctab[c] = reinterpret_cast<const float4 *>(A.ltab)[{4 * lane_off} + c];")
    # one tile: its sub-passes on `tile` and its store
    e("  auto work = [&](float2 *tile, const i64 t) __attribute__((always_inline)) {")
    e("    const i64 base = tile_base(t);")
    e(f"    const i64 block_sample = base >> {n};")
    e(f"    float2 v[{NR}];")
    for si, sub in enumerate(subs):
        _sub_pass(e, sub, si, len(subs), n, R, TB, False, False, measure and si == len(subs) - 1, readout, 0,
                  lanes=lane_k if lanes == "lds" else lanes, interleave=interleave,
                  wave_sync=wave_sync and si + 1 < len(subs) and wave_local(sub, subs[si + 1], TB), swz_mask=30)
    e("    if (A.store) {")
    e("      const int tz = pinv(tid);")
    for it in range(NR // 2):
        e(f"      {{ const i64 gi = base | g_tid | {goff[it]}LL; const int sl = swz_e(2 * ({it * T} + tz));")
        e("        reinterpret_cast<float4 *>(A.state)[gi >> 1] = *reinterpret_cast<const float4 *>(&tile[sl]); }")
    e("    }")
    e("  };")
    # the loads of a tile (and this wave's earlier stores) complete, then every wave's; the buffer
    # refilled next was last read by the store phase before that barrier
    # s_waitcnt vmcnt(0) (expcnt and lgkmcnt unconstrained) through the builtin, so that the
    # compiler's own wait tracking knows the buffer's loads have landed
    wait = "    __builtin_amdgcn_s_waitcnt(0x0F70); __syncthreads();"
    e("  const i64 G = gridDim.x;")
    e("  i64 t = blockIdx.x;")
    e("  if (t < A.ntiles) issue0(t);")
    e("  while (t < A.ntiles) {")
    e(wait)
    e("    if (t + G < A.ntiles) issue1(t + G);")
    e("    work(buf0, t);")
    e("    t += G;")
    e("    if (t >= A.ntiles) break;")
    e(wait)
    e("    if (t + G < A.ntiles) issue0(t + G);")
    e("    work(buf1, t);")
    e("    t += G;")
    e("  }")
    e("}")
    return "\n".join(o)


def pass_source_tail(name: str, words, n: int, R: int, TB: int, lpos, gen: bool, measure: bool, waves: int = 2,
                     readout=None, lanes="lds", lane_off: int = 0, interleave: int = 4,
                     wave_sync: bool = True) -> str:
    """HIP source of one specialised forward pass kernel that streams the next
    tile in during the last sub-pass of the current one.  A resident grid (two
    workgroups per compute unit, one tile buffer each, as pass_source) walks
    tiles; the last sub-pass stores its results from registers straight to the
    state (coalesced runs: the lane bits are the lowest non-register tile bits),
    so once its threads hold their amplitudes the tile buffer is free, and the
    next tile's direct global->LDS loads (``global_load_lds_dwordx4``) run
    while its arithmetic does.  The tile layout keeps amplitude pairs in order
    in one 16-byte slot (``swz_e``).  Tiles within one sample only (n >= R + TB).
    Same sub-pass code, tables and partial layouts as pass_source: results are
    the same bit for bit (measured at 14, 16, 20 and 24 qubits).  Measured no
    faster: 2.38 vs 2.32 ms per 20-qubit batch-64 forward, 3.39 vs 3.14 ms at
    24 qubits x 4."""
    K = R + TB
    T, NR = 1 << TB, 1 << R
    if n < K:
        raise ValueError("tile streaming needs tiles within one sample")
    subs = PL.decode_pass(words, R, TB)
    lpos = [int(x) for x in lpos]
    npos = [p for p in range(n) if p not in lpos]
    nn = len(npos)
    o = []
    e = o.append
    wattr = f" __attribute__((amdgpu_waves_per_eu({waves})))" if waves else ""
    e(f'extern "C" __global__ void __launch_bounds__({T}){wattr}')
    e(f"{name}(JArgs A) {{")
    e(f"  __shared__ __align__(16) float2 tile[{1 << K}];")
    if measure and TB > 6:
        e(f"  __shared__ float red[{(T // 64) * ((K + 1) if readout is None else (n + 1))}];")
    lane_k = None
    if lanes == "lds":
        ents = lane_entries(words, R, TB, False)
        lane_k = {ui: k for k, ui in enumerate(ents)}
        if ents:
            e(f"  __shared__ __align__(16) float4 ctab[{4 * len(ents)}];")
    e("  const int tid = threadIdx.x;")
    e("  auto tile_base = [&](i64 t) -> i64 {")
    e(f"    i64 b = (t >> {nn}) << {n};")
    for i, p in enumerate(npos):
        e(f"    b |= ((t >> {i}) & 1) << {p};")
    e("    return b;")
    e("  };")
    if not gen:
        d_bit, d_inst = async_load_offsets(lpos, R, TB)
        e("  i64 d_tid = 0;")
        for b in range(TB):
            e(f"  d_tid ^= ((tid >> {b}) & 1) ? (i64){d_bit[b]}LL : 0LL;")
        e("  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);")
        e("  auto issue = [&](i64 tt) {")
        e("    const i64 b = tile_base(tt);")
        e("    const float4 *src = reinterpret_cast<const float4 *>(A.src);")
        for ii in range(NR // 2):
            e(f"    __builtin_amdgcn_global_load_lds(src + ((b | (d_tid ^ {d_inst[ii]}LL)) >> 1),")
            e(f"        (__attribute__((address_space(3))) void *)(tile + 2 * ({ii * T} + 64 * wave)), 16, 0, 0);")
        e("  };")
    if lane_k:
        nl = 4 * len(lane_k)
        e(f"  for (int c = tid; c < {nl}; c += {T}) ctab[c] = reinterpret_cast<const float4 *>(A.ltab)[{4 * lane_off} + c];")
    e("  const i64 G = gridDim.x;")
    e("  i64 t = blockIdx.x;")
    if not gen:
        e("  if (t < A.ntiles) issue(t);")
    else:
        e("  __syncthreads();")
    e("  for (; t < A.ntiles; t += G) {")
    if not gen:
        # s_waitcnt vmcnt(0) through the builtin (the compiler's wait tracking sees it): this wave's
        # loads of the tile, and its earlier stores, have completed; then every wave's
        e("    __builtin_amdgcn_s_waitcnt(0x0F70); __syncthreads();")
    e("    const i64 base = tile_base(t);")
    e(f"    const i64 block_sample = base >> {n};")
    e(f"    float2 v[{NR}];")
    prefetch = [] if gen else ["      if (t + G < A.ntiles) issue(t + G);"]
    for si, sub in enumerate(subs):
        last = si == len(subs) - 1
        _sub_pass(e, sub, si, len(subs), n, R, TB, False, gen and si == 0, measure and last, readout, 0,
                  lanes=lane_k if lanes == "lds" else lanes, interleave=interleave,
                  wave_sync=wave_sync and not last and wave_local(sub, subs[si + 1], TB), swz_mask=30,
                  tail=prefetch if last else None)
    if gen:
        e("    __syncthreads();")  # the next tile's first sub-pass writes the tile buffer again
    e("  }")
    e("}")
    return "\n".join(o)


def pass_source_pipelined(name: str, words, n: int, R: int, TB: int, lpos, gen: bool, measure: bool,
                          waves: int = None, readout=None, loaders: int = 8) -> str:
    """HIP source of one specialised forward pass kernel with a load/compute
    pipeline: a persistent workgroup holds two LDS tiles; 2^TB compute
    threads run the sub-passes on one while 2^loaders loader threads store the
    previous tile's result from, and load the next tile into, the other.  The
    HBM traffic of a tile then overlaps the arithmetic of its neighbour inside
    one workgroup.  Every barrier sits outside the role branches, so both
    roles pass the same barriers in the same order.  Tiles of whole samples
    (n < R + TB) are not pipelined."""
    K = R + TB
    T, NR = 1 << TB, 1 << R
    TL = 1 << loaders
    if n < K or loaders > K - 1:
        raise ValueError("pipelined passes need tiles within one sample")
    NL = (1 << (K - 1)) // TL  # float4 slots per loader thread
    subs = PL.decode_pass(words, R, TB)
    lpos = [int(x) for x in lpos]
    npos = [p for p in range(n) if p not in lpos]
    nn = len(npos)
    o = []
    e = o.append
    wattr = f" __attribute__((amdgpu_waves_per_eu({waves})))" if waves else ""
    e(f'extern "C" __global__ void __launch_bounds__({T + TL}){wattr}')
    e(f"{name}(JArgs A) {{")
    e(f"  __shared__ __align__(16) float2 buf[{2 << K}];")
    if measure and TB > 6:
        e(f"  __shared__ float red[{(T // 64) * ((K + 1) if readout is None else (n + 1))}];")
    e(f"  const bool cmp = threadIdx.x < {T};")
    e("  const int tid = cmp ? (int)threadIdx.x : 0;")
    e(f"  const int lt = cmp ? 0 : (int)threadIdx.x - {T};")
    e(f"  const i64 total = A.B << {n};")
    e("  i64 g_ld = 0;")
    for i in range(loaders):
        e(f"  g_ld |= (i64)((lt >> {i}) & 1) << {lpos[i + 1]};")
    loff = [sum(1 << lpos[loaders + 1 + i] for i in range(K - 1 - loaders) if (j >> i) & 1) for j in range(NL)]
    e("  auto tile_base = [&](i64 t) -> i64 {")
    e(f"    i64 b = (t >> {nn}) << {n};")
    for i, p in enumerate(npos):
        e(f"    b |= ((t >> {i}) & 1) << {p};")
    e("    return b;")
    e("  };")
    e(f"  float4 ld[{NL}];")
    e("  auto fetch = [&](i64 b) {")
    for j in range(NL):
        e(f"    ld[{j}] = reinterpret_cast<const float4 *>(A.src)[(b | g_ld | {loff[j]}LL) >> 1];")
    e("  };")
    e("  auto to_lds = [&](int bo) {")
    e("    const int lz = pinv(lt);")
    for j in range(NL):
        e(f"    {{ const int sl = swz(2 * ({j * TL} + lz)); const float4 val = ld[{j}]; const bool sw = sl & 1;")
        e("      *reinterpret_cast<float4 *>(&buf[bo + (sl & ~1)]) = make_float4(sw ? val.z : val.x, sw ? val.w : val.y,"
          " sw ? val.x : val.z, sw ? val.y : val.w); }")
    e("  };")
    e("  auto store_tile = [&](i64 b, int bo) {")
    e("    const int lz = pinv(lt);")
    for j in range(NL):
        e(f"    {{ const int sl = swz(2 * ({j * TL} + lz));")
        e("      const float4 raw = *reinterpret_cast<const float4 *>(&buf[bo + (sl & ~1)]); const bool sw = sl & 1;")
        e(f"      reinterpret_cast<float4 *>(A.state)[(b | g_ld | {loff[j]}LL) >> 1] = make_float4(sw ? raw.z : raw.x,"
          " sw ? raw.w : raw.y, sw ? raw.x : raw.z, sw ? raw.y : raw.w); }")
    e("  };")
    e("  const i64 G = gridDim.x;")
    e("  i64 t = blockIdx.x;")
    if not gen:
        e("  if (!cmp && t < A.ntiles) { fetch(tile_base(t)); to_lds(0); }")
    e("  __syncthreads();")
    e("  int k = 0;")
    e("  for (; t < A.ntiles; ++k, t += G) {")
    e(f"    const int boff = (k & 1) << {K};")
    e("    const i64 base = tile_base(t);")
    e(f"    const i64 block_sample = base >> {n};")
    e("    if (!cmp) {")
    e(f"      if (k > 0 && A.store) store_tile(tile_base(t - G), boff ^ {1 << K});")
    if not gen:
        e("      if (t + G < A.ntiles) fetch(tile_base(t + G));")
    e("    }")
    e(f"    float2 v[{NR}];")
    for si, sub in enumerate(subs):
        pipe = {}
        e("    if (cmp) {")
        _sub_pass(e, sub, si, len(subs), n, R, TB, False, gen and si == 0, measure and si == len(subs) - 1, readout,
                  0, pipe)
        e("    }")
        e("    __syncthreads();")
        if "finish" in pipe:
            e("    if (cmp) {")
            for line in pipe["finish"]:
                e(line)
            e("    }")
    if not gen:
        e(f"    if (!cmp && t + G < A.ntiles) to_lds(boff ^ {1 << K});")
    e("    __syncthreads();")
    e("  }")
    e(f"  if (!cmp && k > 0 && A.store) store_tile(tile_base(t - G), ((k - 1) & 1) << {K});")
    e("}")
    return "\n".join(o)


def wave_local(a, b, TB: int) -> bool:
    """True when sub-passes a and b (decoded) map the wave-index bits (thread
    bits 6 and up) to the same tile bits: the tile elements a wave holds are
    then the same set in both, so passing them through LDS needs no block barrier."""
    return list(a.G[6:TB]) == list(b.G[6:TB])


def _sub_pass(e, sub, s, nsub, n, R, TB, multi, gen, measure, readout=None, valu_pairs=0, pipe=None, lanes=False,
              interleave=False, wave_sync=False, swz_mask=31, tail=None, direct_load=False):
    """One sub-pass.  ``pipe`` (pipelined kernels): a dict; the tile is
    ``buf[boff + i]``, no closing barrier is emitted, and a measurement's
    cross-wave finish is left in ``pipe["finish"]`` for after the caller's
    barrier.  ``tail`` (the last sub-pass of a kernel that stores from
    registers): the results go straight from registers to the state instead of
    back into the tile; a non-empty list of lines is emitted after a block
    barrier that follows the register loads (the tile buffer is then free).
    ``direct_load`` (the first sub-pass of a kernel that loads into registers):
    the amplitudes come straight from the state (``A.src``) instead of the tile."""
    if pipe is not None:
        import re

        inner = e
        sub_tile = re.compile(r"tile\[([^\]]*)\]")

        def e(line):  # noqa: F811
            inner(sub_tile.sub(r"buf[boff + (\1)]", line))
    K = R + TB
    T, NR = 1 << TB, 1 << R
    e(f"    {{  // sub-pass {s}")
    e("      const int tq = pinv(tid);")
    e("      int gl = 0; i64 gbase = base;")
    for i in range(TB):
        e(f"      gl |= ((tq >> {i}) & 1) << {sub.G[i]}; gbase |= (i64)((tq >> {i}) & 1) << {sub.GG[i]};")
    e(f"      const i64 sample = {'gbase >> ' + str(n) if multi else 'block_sample'};")
    e("      const bool live = sample < A.B;")
    e("      const i64 csample = live ? sample : 0;")
    e(f"      const int lb = {'swz' if swz_mask == 31 else 'swz_e'}(gl);")
    if pipe is None:
        e("      const int lb8 = lb << 3;")
    sw = [_swz(1 << sub.P[r], swz_mask) for r in range(R)]
    swj = []
    for j in range(NR):
        a = 0
        for r in range(R):
            if (j >> r) & 1:
                a ^= sw[r]
        swj.append(a)
    cq = {}
    if isinstance(lanes, dict):
        # the sub-pass's operands are read from LDS up front (they are small and ready before the tile)
        for si_, (bits, _c) in enumerate(sub.segments):
            for r, ui, ps in bits:
                if not ps and ui in lanes:
                    nm = f"cq{si_}_{r}"
                    cq[(si_, r)] = nm
                    e(f"      const float4 {nm} = ctab[{4 * lanes[ui]} + (tq & 3)];")
    elif lanes:
        e("      const float *lt = A.ltab + 4 * (tq & 3);  // tq is pinned per sub-pass: the loads stay here")
        for si_, (bits, _c) in enumerate(sub.segments):
            for r, ui, ps in bits:
                if ps and multi:
                    continue
                idx = f"({ui} + csample)" if ps else f"{ui}"
                nm = f"cq{si_}_{r}"
                cq[(si_, r)] = nm
                e(f"      const float4 {nm} = *reinterpret_cast<const float4 *>(lt + 16 * {idx});")
    if gen:
        rmask = 0
        for r in range(R):
            rmask |= 1 << sub.GP[r]
        e("      if (live) {")
        e(f"        const float2 *E = reinterpret_cast<const float2 *>(A.efac) + sample * {2 * n};")
        e("        float2 pf = make_float2(1.f, 0.f);")
        for pos in range(n):
            if not (rmask >> pos) & 1:
                e(f"        pf = cmul(pf, E[{2 * pos} + ((gbase >> {pos}) & 1)]);")
        e("        v[0] = pf;")
        for r in range(R):
            e(f"        {{ const float2 f0 = E[{2 * sub.GP[r]}], f1 = E[{2 * sub.GP[r] + 1}];")
            for j in range(1 << r):
                e(f"          v[{j | (1 << r)}] = cmul(v[{j}], f1); v[{j}] = cmul(v[{j}], f0);")
            e("        }")
        e("      } else {")
        for j in range(NR):
            e(f"        v[{j}] = make_float2(0.f, 0.f);")
        e("      }")
    elif direct_load:
        _direct_load(e, sub, R)
    else:
        for j in range(NR):
            if pipe is None:
                e(f"      v[{j}] = lds_at(tile, lb8 ^ {8 * swj[j]});")
            else:
                e(f"      v[{j}] = tile[lb ^ {swj[j]}];")
    if tail:
        e("      __syncthreads();  // every thread holds its amplitudes: the tile buffer is free")
        for line in tail:
            e(line)
    phys = list(range(NR))
    for si_, (bits, cnots) in enumerate(sub.segments):
        for r, ui, ps in bits:
            off = f"8 * ({ui} + csample)" if ps else f"{8 * ui}"
            e("      {")
            if (si_, r) in cq:
                nm = cq[(si_, r)]
                e(f"        const float m[4] = {{{nm}.x, {nm}.y, {nm}.z, {nm}.w}};")
                prs = [(phys[j], phys[j | (1 << r)]) for j in range(NR) if not (j >> r) & 1]
                # the last `valu_pairs` pairs run on the vector ALU, spread between the
                # matrix-core groups so that both units have work at the same time
                nv = min(int(valu_pairs), len(prs))
                vprs = prs[len(prs) - nv:]
                prs = prs[:len(prs) - nv]
                if vprs:
                    e("        const int *ut = pin(A.utab);")
                step = 4 if interleave == 4 else (2 if interleave else 1)
                q = 0
                while q < len(prs):
                    grp = prs[q:q + step]
                    if len(grp) == 4:
                        e("        mfma_pair4(" + ", ".join(f"v[{a}], v[{b}]" for a, b in grp) + ", m);")
                    elif len(grp) >= 2:
                        (a0, b0), (a1, b1) = grp[0], grp[1]
                        e(f"        mfma_pair2(v[{a0}], v[{b0}], v[{a1}], v[{b1}], m);")
                        grp = grp[:2]
                    else:
                        e(f"        mfma_pair(v[{grp[0][0]}], v[{grp[0][1]}], m);")
                    q += len(grp)
                    if vprs:
                        a, b = vprs.pop(0)
                        e(f"        valu_pair_s(v[{a}], v[{b}], ut + {off});")
                for a, b in vprs:
                    e(f"        valu_pair_s(v[{a}], v[{b}], ut + {off});")
            elif ps and multi:
                e("        const int *ut = pin(A.utab);")
                e(f"        const int *u = ut + {off};")
                for j in range(NR):
                    if not (j >> r) & 1:
                        e(f"        valu_pair(v[{phys[j]}], v[{phys[j | (1 << r)]}], u);")
            else:
                e("        const int *ut = pin(A.utab);")
                e(f"        float m[4]; coeffs(ut + {off}, m);")
                k = 0
                for j in range(NR):
                    if not (j >> r) & 1:
                        if k < valu_pairs:
                            e(f"        valu_pair_s(v[{phys[j]}], v[{phys[j | (1 << r)]}], ut + {off});")
                        else:
                            e(f"        mfma_pair(v[{phys[j]}], v[{phys[j | (1 << r)]}], m);")
                        k += 1
            e("      }")
        for uniform, tgt, ctl in cnots:
            if not uniform:
                for j in range(NR):
                    if (j >> ctl) & 1 and not (j >> tgt) & 1:
                        k = j | (1 << tgt)
                        phys[j], phys[k] = phys[k], phys[j]
            else:
                e(f"      {{ const bool cnd = (gbase >> {ctl}) & 1;")
                for j in range(NR):
                    if not (j >> tgt) & 1:
                        a, b = phys[j], phys[j | (1 << tgt)]
                        e(f"        {{ const float2 x = v[{a}], y = v[{b}]; v[{a}] = cnd ? y : x; v[{b}] = cnd ? x : y; }}")
                e("      }")
    if sub.phases:
        e("      {")
        e("        float pr = 1.f, pi = 0.f;")
        for pos, ci, fl in sub.phases:
            if fl & PL.SEG_NOUNAPPLY:
                continue
            idx = f"({ci} + csample)" if fl & PL.SEG_PER_SAMPLE else f"{ci}"
            e(f"        {{ const float c = __int_as_float(as_const(A.cs)[2 * {idx}]), "
              f"sn0 = __int_as_float(as_const(A.cs)[2 * {idx} + 1]);")
            e(f"          const float sn = ((gbase >> {pos}) & 1) ? -sn0 : sn0;")
            e("          const float nr = pr * c + pi * sn, ni = pi * c - pr * sn; pr = nr; pi = ni; }")
        e("        const float2 z = make_float2(pr, pi);")
        for j in range(NR):
            e(f"        v[{j}] = cmul(z, v[{j}]);")
        e("      }")
    if tail is not None:
        _direct_store(e, sub, phys, R)
        if measure:
            _measure(e, sub, phys, n, R, TB, multi, readout, pipe)
        e("    }")
        return
    for j in range(NR):
        if pipe is None:
            e(f"      lds_at(tile, lb8 ^ {8 * swj[j]}) = v[{phys[j]}];")
        else:
            e(f"      tile[lb ^ {swj[j]}] = v[{phys[j]}];")
    if measure:  # after the tile writes: the amplitudes' registers are free for the partials
        _measure(e, sub, phys, n, R, TB, multi, readout, pipe)
    if pipe is None and wave_sync:
        # the next sub-pass keeps this one's wave-index bits: every wave reads back only the
        # quarter of the tile it wrote, and a wave's LDS operations complete in order
        e("      __builtin_amdgcn_fence(__ATOMIC_RELEASE, \"wavefront\"); __builtin_amdgcn_wave_barrier();")
        e("      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, \"wavefront\");")
    elif pipe is None:
        e("      __syncthreads();")
    e("    }")


def _register_offsets(sub, R):
    """Amplitude offset (from the thread's gbase) of each register of a sub-pass, and
    the register bit at amplitude bit 0 (None if none is)."""
    roff = [sum(1 << sub.GP[r] for r in range(R) if (j >> r) & 1) for j in range(1 << R)]
    return roff, next((r for r in range(R) if sub.GP[r] == 0), None)


def _full_lines(sub, R) -> bool:
    """Whether one wave's direct register loads or stores of a sub-pass fill whole
    128-byte lines: the lowest lane bits sit on amplitude bits 0..3 (8-byte accesses)
    or, with a register on amplitude bit 0, on bits 1..3 (16-byte accesses)."""
    _roff, r0 = _register_offsets(sub, R)
    low = [1, 2, 3] if r0 is not None else [0, 1, 2, 3]
    return [sub.GG[b] for b in range(len(low))] == low


def _direct_load(e, sub, R):
    """Load a sub-pass's registers straight from the state: register j of a thread
    is amplitude gbase | sum of 1 << GP[r] over its set bits r (the lane bits are the
    lowest non-register tile bits, so a wave's loads form contiguous runs; with a
    register bit at amplitude bit 0, two registers come from one 16-byte load)."""
    NR = 1 << R
    roff, r0 = _register_offsets(sub, R)
    e("      {")
    e("        const float4 *i4 = reinterpret_cast<const float4 *>(A.src);")
    e("        const float2 *i2 = reinterpret_cast<const float2 *>(A.src);")
    for j in range(NR):
        if r0 is not None:
            if (j >> r0) & 1:
                continue
            e(f"        {{ const float4 q = i4[(gbase | {roff[j]}LL) >> 1]; v[{j}] = make_float2(q.x, q.y); "
              f"v[{j | (1 << r0)}] = make_float2(q.z, q.w); }}")
        else:
            e(f"        v[{j}] = i2[gbase | {roff[j]}LL];")
    e("      }")


def _direct_store(e, sub, phys, R):
    """Store a sub-pass's registers straight to the state: register j of a thread
    is amplitude gbase | sum of 1 << GP[r] over its set bits r.  The lane bits hold
    the lowest non-register tile bits in ascending order (planner.thread_bits), so
    a wave's stores form contiguous runs; when a register bit sits at amplitude
    bit 0, two registers make one 16-byte pair store."""
    NR = 1 << R
    roff, r0 = _register_offsets(sub, R)
    e("      if (A.store) {")
    e("        float4 *o4 = reinterpret_cast<float4 *>(A.state);")
    e("        float2 *o2 = reinterpret_cast<float2 *>(A.state);")
    for j in range(NR):
        if r0 is not None:
            if (j >> r0) & 1:
                continue
            a, b = phys[j], phys[j | (1 << r0)]
            e(f"        o4[(gbase | {roff[j]}LL) >> 1] = make_float4(v[{a}].x, v[{a}].y, v[{b}].x, v[{b}].y);")
        else:
            e(f"        o2[gbase | {roff[j]}LL] = v[{phys[j]}];")
    e("      }")


def _parity(var: str, mask: int) -> str:
    """Expression of the parity of ``var & mask`` (a 64-bit index)."""
    if mask & (mask - 1) == 0:
        return f"(({var} >> {mask.bit_length() - 1}) & 1)"
    return f"(__popcll({var} & {mask}LL) & 1)"


def _measure(e, sub, phys, n, R, TB, multi, readout=None, pipe=None):
    """MeasureProbability partials of the tile: K + 1 per slot by local bit
    (layout of qs_fused_measure_finish) or, with ``readout`` masks, n + 1 per
    slot by bit position (layout of qs_readout_finish)."""
    K = R + TB
    T, NR = 1 << TB, 1 << R
    e("      {")
    e(f"        float p[{NR}]; float tot = 0.f;")
    for j in range(NR):
        e(f"        p[{j}] = v[{phys[j]}].x * v[{phys[j]}].x + v[{phys[j]}].y * v[{phys[j]}].y; tot += p[{j}];")
    if readout is None:
        W = K + 1
        e(f"        float c[{W}];")
        for r in range(R):
            terms = " + ".join(f"p[{j}]" for j in range(NR) if not (j >> r) & 1)
            e(f"        c[{r}] = {terms};")
        for i in range(TB):
            e(f"        c[{R + i}] = ((tid >> {i}) & 1) ? 0.f : tot;")
        e(f"        c[{K}] = tot;")
        dst = [sub.P[q] if q < R else (sub.G[q - R] if q < K else K) for q in range(W)]
    else:
        # measured bit p = parity(a_p & x): the register part of a_p selects
        # which of the thread's amplitudes count, the rest (thread and tile
        # bits) flips the thread's share to its complement
        W = n + 1
        regpos = 0
        for r in range(R):
            regpos |= 1 << sub.GP[r]
        e(f"        float c[{W}];")
        sums = {}
        for pos in range(n):
            a = int(readout[pos])
            mr = sum(1 << r for r in range(R) if (a >> sub.GP[r]) & 1)
            rest = a & ~regpos
            if mr and mr not in sums:
                terms = " + ".join(f"p[{j}]" for j in range(NR) if bin(j & mr).count("1") % 2 == 0)
                sums[mr] = f"s{mr}"
                e(f"        const float s{mr} = {terms};")
            s0 = sums[mr] if mr else "tot"
            if rest:
                e(f"        c[{pos}] = {_parity('gbase', rest)} ? tot - {s0} : {s0};")
            else:
                e(f"        c[{pos}] = {s0};")
        e(f"        c[{n}] = tot;")
        dst = list(range(W))
    seglog = (n - R) if multi else TB
    wlog = min(seglog, 6)
    red = _slot_sum("c[q]", wlog)  # DPP lane sums (no LDS traffic) where the slot width allows
    if red is not None:
        e(f"        for (int q = 0; q < {W}; ++q) c[q] = {red};")
    else:
        e(f"        for (int q = 0; q < {W}; ++q) {{ float x = c[q];")
        e(f"          for (int off = 1; off < {1 << wlog}; off <<= 1) x += __shfl_xor(x, off, 64); c[q] = x; }}")
    spt_log = (K - n) if multi else 0
    e(f"        float *mp = A.mpart + (t << {spt_log}) * {W};")
    if seglog <= 6:
        e(f"        if ((tid & {(1 << seglog) - 1}) == 0) {{")
        e(f"          float *ob = mp + (tid >> {seglog}) * {W};")
        for q in range(W):
            e(f"          ob[{dst[q]}] = c[{q}];")
        e("        }")
    else:
        wps = 1 << (seglog - 6)
        nsamp = (T // 64) // wps
        e("        if ((tid & 63) == 0) {")
        e(f"          for (int q = 0; q < {W}; ++q) red[(tid >> 6) * {W} + q] = c[q];")
        e("        }")
        fin = [f"        if (tid < {nsamp * W}) {{",
               f"          const int sm = tid / {W}, q = tid % {W};",
               "          float x = 0.f;",
               f"          for (int k = 0; k < {wps}; ++k) x += red[(sm * {wps} + k) * {W} + q];",
               f"          const int dst[{W}] = {{{', '.join(str(d) for d in dst)}}};",
               f"          A.mpart[((t << {spt_log}) + sm) * {W} + dst[q]] = x;",
               "        }"]
        if pipe is not None:
            pipe["finish"] = fin
        else:
            e("        __syncthreads();")
            for line in fin:
                e(line)
    e("      }")


def adjoint_lane_entries(words, R: int, TB: int):
    """utab indices of the uniform unitaries an adjoint step un-applies,
    distinct, in program order (its block of the U^dagger lane table)."""
    out = []
    for sub in PL.decode_adjoint_pass(words, R, TB):
        for _cnots, bits in sub.segments:
            for _r, u, ps, probe, _q, _rows in bits:
                if not ps and not probe and u not in out:
                    out.append(int(u))
    return out


def adjoint_source(name: str, words, n: int, R: int, TB: int, lpos, seed: bool,
                   prefetch: bool = False, waves: int = None, readout=None, lanes=False,
                   split_pairs: bool = False, lane_off: int = 0, interleave: bool = False,
                   wave_sync: bool = False, hermitian: bool = False, direct_load: bool = False,
                   direct_store: bool = False, rows4: bool = True) -> str:
    """HIP source of one specialised adjoint pass kernel (the adjoint tile of
    qs_adjoint_pass with the pass program compiled in).  ``seed``: the first
    step of the sweep, whose co-state is built from psi and w = dL/dP, with
    the measured bit at position p the parity of ``readout[p]`` (None: bit p).
    ``lanes`` = "lds" and ``split_pairs`` as for pass_source, with the U^dagger
    lane table (entries ``lane_off`` onward, adjoint_lane_entries order);
    ``interleave`` and ``wave_sync`` as for pass_source.  ``hermitian``: the
    gradient of every fused rotation reads Im Tr(Q M) through the four real
    combinations of M that a Hermitian generator Q weighs (herm_trace), 12
    fused multiply-adds per amplitude pair instead of 16 (and of 24 separate
    products and sums); equal to float rounding.  ``direct_load`` /
    ``direct_store``: the first sub-pass reads psi and lambda straight from the
    state and the last writes them straight back, instead of staging the tiles
    through LDS (tiles within one sample; direct_load not for the seeding step,
    which builds lambda while staging, nor with prefetch); same results bit for
    bit.  ``rows4``: a unitary's gradient rows are summed over the wave four at
    a time (wave_sum4: two exchange rounds, row rotations and the row and half
    swaps, the sums landing in lanes 0-3 for one store) instead of one DPP and
    readlane reduction per row: the same additions in the same tree, so the
    gradients are bit for bit the same (4.09-4.17 vs 4.17-4.22 ms for the 15
    steps at 16 qubits x 256 x 20 layers)."""
    K = R + TB
    T, NR = 1 << TB, 1 << R
    multi = n < K
    direct_load = direct_load and not multi and not seed and not prefetch
    direct_store = direct_store and not multi
    subs = PL.decode_adjoint_pass(words, R, TB)
    lpos = [int(x) for x in lpos]
    npos = [] if multi else [p for p in range(n) if p not in lpos]
    nn = len(npos)
    seglog = (n - R) if multi else TB
    wlog = min(seglog, 6)
    o = []
    e = o.append
    wattr = f" __attribute__((amdgpu_waves_per_eu({waves})))" if waves else ""
    e(f'extern "C" __global__ void __launch_bounds__({T}){wattr}')
    e(f"{name}(JAdj A) {{")
    e(f"  __shared__ __align__(16) float2 tp[{1 << K}];")
    e(f"  __shared__ __align__(16) float2 tl[{1 << K}];")
    lane_k = None
    if lanes == "lds":
        ents = adjoint_lane_entries(words, R, TB)
        lane_k = {u: k for k, u in enumerate(ents)}
        if ents:
            e(f"  __shared__ __align__(16) float4 ctab[{4 * len(ents)}];")
    e("  const int tid = threadIdx.x;")
    e(f"  const i64 total = A.B << {n};")
    e("  i64 g_tid = 0;")
    for i in range(TB):
        e(f"  g_tid |= (i64)((tid >> {i}) & 1) << {lpos[i + 1]};")
    goff = [sum(1 << lpos[TB + 1 + i] for i in range(R - 1) if (it >> i) & 1) for it in range(NR // 2)]
    e(f"  float4 pa[{max(NR // 2, 1)}], pb[{max(NR // 2, 1)}];")
    e("  auto tile_base = [&](i64 t) -> i64 {")
    if multi:
        e(f"    return t << {K};")
    else:
        e(f"    i64 b = (t >> {nn}) << {n};")
        for i, p in enumerate(npos):
            e(f"    b |= ((t >> {i}) & 1) << {p};")
        e("    return b;")
    e("  };")
    e("  auto fetch = [&](i64 base) {")
    for it in range(NR // 2):
        e(f"    {{ const i64 gi = base | g_tid | {goff[it]}LL;")
        ld_a = "reinterpret_cast<const float4 *>(A.psi_src)[gi >> 1]"
        ld_b = "reinterpret_cast<const float4 *>(A.lam)[gi >> 1]"
        if multi:
            e(f"      pa[{it}] = gi < total ? {ld_a} : make_float4(0.f, 0.f, 0.f, 0.f);")
            if not seed:
                e(f"      pb[{it}] = gi < total ? {ld_b} : make_float4(0.f, 0.f, 0.f, 0.f);")
        else:
            e(f"      pa[{it}] = {ld_a};")
            if not seed:
                e(f"      pb[{it}] = {ld_b};")
        e("    }")
    e("  };")
    e("  i64 t = blockIdx.x;")
    if prefetch:
        e("  if (t < A.ntiles) fetch(tile_base(t));")
    e("  for (; t < A.ntiles; t += gridDim.x) {")
    e("    const i64 base = tile_base(t);")
    if not prefetch and not direct_load:
        e("    fetch(base);")
    e(f"    const i64 block_sample = base >> {n};")
    e("    const int ts = pinv(tid);")
    if lane_k:
        nl = 4 * len(lane_k)
        e(f"    for (int c = ts; c < {nl}; c += {T}) ctab[c] = reinterpret_cast<const float4 *>(A.ltab)[{4 * lane_off} + c];")
    for it in range(0 if direct_load else NR // 2):
        e(f"    {{ const int sl = swz(2 * ({it * T} + ts)); const bool sw = sl & 1; const float4 a = pa[{it}];")
        if seed:
            # co-state of MeasureProbability for the pair (gi, gi + 1)
            e(f"      const i64 gi = base | g_tid | {goff[it]}LL;")
            e(f"      const i64 smp = {'gi >> ' + str(n) if multi else 'block_sample'};")
            e(f"      const float *wb = A.seed_w + (gi < total ? smp : 0) * {n};")
            # gi is even: wc collects the weights whose measured bit is 0 at both gi
            # and gi + 1, wa (wz) those measured 0 only at gi (gi + 1)
            e("      float wc = 0.f, wa = 0.f, wz = 0.f;")
            masks = [1 << p for p in range(n)] if readout is None else [int(x) for x in readout]
            for p in list(range(1, n)) + [0]:
                a = masks[p]
                if not a & 1:
                    e(f"      wc += {_parity('gi', a)} ? 0.f : wb[{n - 1 - p}];")
                elif a == 1:
                    e(f"      wa += wb[{n - 1 - p}];")
                else:
                    e(f"      {{ const bool o = {_parity('gi', a ^ 1)}; wa += o ? 0.f : wb[{n - 1 - p}]; "
                      f"wz += o ? wb[{n - 1 - p}] : 0.f; }}")
            e("      const float w0 = 2.f * (wc + wa), w1 = 2.f * (wc + wz);")
            e("      const float4 b = make_float4(w0 * a.x, w0 * a.y, w1 * a.z, w1 * a.w);")
        else:
            e(f"      const float4 b = pb[{it}];")
        if split_pairs:
            e("      (void)sw; tp[sl] = make_float2(a.x, a.y); tp[sl ^ 1] = make_float2(a.z, a.w);")
            e("      tl[sl] = make_float2(b.x, b.y); tl[sl ^ 1] = make_float2(b.z, b.w); }")
        else:
            e("      *reinterpret_cast<float4 *>(&tp[sl & ~1]) = make_float4(sw ? a.z : a.x, sw ? a.w : a.y, sw ? a.x : a.z,"
              " sw ? a.y : a.w);")
            e("      *reinterpret_cast<float4 *>(&tl[sl & ~1]) = make_float4(sw ? b.z : b.x, sw ? b.w : b.y, sw ? b.x : b.z,"
              " sw ? b.y : b.w); }")
    e("    __syncthreads();")
    if prefetch:
        e("    { const i64 tn = t + gridDim.x; if (tn < A.ntiles) fetch(tile_base(tn)); }")
    e(f"    const i64 uslot = t * {T >> wlog} + (tid >> {wlog});")
    e(f"    const bool leader = (tid & {(1 << wlog) - 1}) == 0;")
    e(f"    float2 vp[{NR}], vl[{NR}];")
    for s, sub in enumerate(subs):
        _adj_sub_pass(e, sub, s, n, R, TB, multi, wlog, lane_k, interleave,
                      wave_sync and s + 1 < len(subs) and wave_local(sub, subs[s + 1], TB), hermitian,
                      direct_load=direct_load and s == 0, direct_store=direct_store and s == len(subs) - 1,
                      rows4=rows4)
    if direct_store:
        e("    __syncthreads();")
        e("  }")
        e("}")
        return "\n".join(o)
    e("    const int tz = pinv(tid);")
    for it in range(NR // 2):
        e(f"    {{ const i64 gi = base | g_tid | {goff[it]}LL; const int sl = swz(2 * ({it * T} + tz)); const bool sw = sl & 1;")
        cond = "if (gi < total) " if multi else ""
        if split_pairs:
            e("      (void)sw; const float2 a0 = tp[sl], a1 = tp[sl ^ 1], b0 = tl[sl], b1 = tl[sl ^ 1];")
            e(f"      {cond}{{ reinterpret_cast<float4 *>(A.psi)[gi >> 1] = make_float4(a0.x, a0.y, a1.x, a1.y);")
            e("        reinterpret_cast<float4 *>(A.lam)[gi >> 1] = make_float4(b0.x, b0.y, b1.x, b1.y); } }")
        else:
            e("      const float4 a = *reinterpret_cast<const float4 *>(&tp[sl & ~1]);")
            e("      const float4 b = *reinterpret_cast<const float4 *>(&tl[sl & ~1]);")
            e(f"      {cond}{{ reinterpret_cast<float4 *>(A.psi)[gi >> 1] = make_float4(sw ? a.z : a.x, sw ? a.w : a.y,"
              " sw ? a.x : a.z, sw ? a.y : a.w);")
            e("        reinterpret_cast<float4 *>(A.lam)[gi >> 1] = make_float4(sw ? b.z : b.x, sw ? b.w : b.y,"
              " sw ? b.x : b.z, sw ? b.y : b.w); } }")
    e("    __syncthreads();")
    e("  }")
    e("}")
    return "\n".join(o)


def _slot_sum(expr: str, wlog: int) -> str:
    """Expression summing ``expr`` over the 2^wlog lanes of a reduction slot
    (valid in the slot's first lane)."""
    if wlog == 6:
        return f"wave_sum({expr})"
    if wlog == 5:
        return f"half_sum({expr})"
    if wlog == 4:
        return f"row_sum({expr})"
    if wlog == 2:
        return f"quad_sum({expr})"
    if wlog == 1:
        return f"dpp_add<0xB1>({expr})"
    if wlog == 0:
        return expr
    return None


def _emit_put(e, expr: str, row: int, wlog: int, indent: str):
    """Reduce ``expr`` over the slot's lanes and store it as gradient partial ``row``."""
    red = _slot_sum(expr, wlog)
    if red is not None:
        e(f"{indent}{{ const float g = {red};")
    else:
        e(f"{indent}{{ float g = {expr};")
        e(f"{indent}  for (int off = 1; off < {1 << wlog}; off <<= 1) g += __shfl_xor(g, off, 64);")
    e(f"{indent}  if (leader) pinf(A.gpart)[(i64){row} * A.NS + uslot] = g; }}")


def _adj_sub_pass(e, sub, s, n, R, TB, multi, wlog, lanes=None, interleave=False, wave_sync=False,
                  hermitian=False, direct_load=False, direct_store=False, rows4=True):
    """One adjoint sub-pass.  ``direct_load`` (the first sub-pass): psi and lambda
    come straight from the state (psi from ``A.psi_src``) instead of the tiles;
    ``direct_store`` (the last): they go straight back to the state."""
    NR = 1 << R
    e(f"    {{  // adjoint sub-pass {s}")
    e("      const int tq = pinv(tid);")
    e("      int gl = 0; i64 gbase = base;")
    for i in range(TB):
        e(f"      gl |= ((tq >> {i}) & 1) << {sub.G[i]}; gbase |= (i64)((tq >> {i}) & 1) << {sub.GG[i]};")
    e(f"      const i64 sample = {'gbase >> ' + str(n) if multi else 'block_sample'};")
    e("      const i64 csample = sample < A.B ? sample : 0;")
    e("      const int lb = swz(gl);")
    e("      const int lb8 = lb << 3;")
    sw = [_swz(1 << sub.P[r]) for r in range(R)]
    swj = []
    for j in range(NR):
        a = 0
        for r in range(R):
            if (j >> r) & 1:
                a ^= sw[r]
        swj.append(a)
    if direct_load:
        roff, r0 = _register_offsets(sub, R)
        e("      {")
        for j in range(NR):
            if r0 is not None:
                if (j >> r0) & 1:
                    continue
                k = j | (1 << r0)
                e(f"        {{ const i64 q = (gbase | {roff[j]}LL) >> 1;")
                e("          const float4 a = reinterpret_cast<const float4 *>(A.psi_src)[q], "
                  "b = reinterpret_cast<const float4 *>(A.lam)[q];")
                e(f"          vp[{j}] = make_float2(a.x, a.y); vp[{k}] = make_float2(a.z, a.w); "
                  f"vl[{j}] = make_float2(b.x, b.y); vl[{k}] = make_float2(b.z, b.w); }}")
            else:
                e(f"        vp[{j}] = reinterpret_cast<const float2 *>(A.psi_src)[gbase | {roff[j]}LL]; "
                  f"vl[{j}] = reinterpret_cast<const float2 *>(A.lam)[gbase | {roff[j]}LL];")
        e("      }")
    else:
        for j in range(NR):
            e(f"      vp[{j}] = lds_at(tp, lb8 ^ {8 * swj[j]}); vl[{j}] = lds_at(tl, lb8 ^ {8 * swj[j]});")
    cq = {}
    if lanes:
        for si_, (_cn, bits) in enumerate(sub.segments):
            for r, u, ps, probe, _q, _rows in bits:
                if not ps and not probe and u in lanes:
                    nm = f"cq{si_}_{r}"
                    cq[(si_, r)] = nm
                    e(f"      const float4 {nm} = ctab[{4 * lanes[u]} + (tq & 3)];")

    phys = list(range(NR))
    if sub.phases:
        terms = " + ".join(f"(vl[{j}].x * vp[{j}].y - vl[{j}].y * vp[{j}].x)" for j in range(NR))
        e(f"      {{ const float S = {terms};")
        for pos, ci, fl, row in sub.phases:
            if row >= 0:
                _emit_put(e, f"(((gbase >> {pos}) & 1) ? -S : S)", row, wlog, "        ")
        e("        float pr = 1.f, pi = 0.f;")
        for pos, ci, fl, row in sub.phases:
            if fl & PL.SEG_NOUNAPPLY:
                continue
            idx = f"({ci} + csample)" if fl & PL.SEG_PER_SAMPLE else f"{ci}"
            e(f"        {{ const float c = __int_as_float(as_const(A.cs)[2 * {idx}]), "
              f"sn0 = __int_as_float(as_const(A.cs)[2 * {idx} + 1]);")
            # inverse phase: the sign flips
            e(f"          const float sn = ((gbase >> {pos}) & 1) ? sn0 : -sn0;")
            e("          const float nr = pr * c + pi * sn, ni = pi * c - pr * sn; pr = nr; pi = ni; }")
        e("        const float2 z = make_float2(pr, pi);")
        for j in range(NR):
            e(f"        vp[{j}] = cmul(z, vp[{j}]); vl[{j}] = cmul(z, vl[{j}]);")
        e("      }")
    for si_, (cnots, bits) in enumerate(sub.segments):
        for uniform, tgt, ctl in cnots:
            if not uniform:
                for j in range(NR):
                    if (j >> ctl) & 1 and not (j >> tgt) & 1:
                        k = j | (1 << tgt)
                        phys[j], phys[k] = phys[k], phys[j]
            else:
                e(f"      {{ const bool cnd = (gbase >> {ctl}) & 1;")
                for j in range(NR):
                    if not (j >> tgt) & 1:
                        a, b = phys[j], phys[j | (1 << tgt)]
                        e(f"        {{ const float2 x = vp[{a}], y = vp[{b}]; vp[{a}] = cnd ? y : x; vp[{b}] = cnd ? x : y; }}")
                        e(f"        {{ const float2 x = vl[{a}], y = vl[{b}]; vl[{a}] = cnd ? y : x; vl[{b}] = cnd ? x : y; }}")
                e("      }")
        for r, u, ps, probe, q, rows in bits:
            e("      {")
            e("        const int *ut = pin(A.utab), *qt = pin(A.qtab);")
            if hermitian:
                # M = sum of psi-pair x lambda-pair^dagger enters only through Im Tr(Q M) with Q
                # Hermitian: four fused multiply-add chains instead of the eight entries of M
                e("        float h[4] = {0.f, 0.f, 0.f, 0.f};")
                for j in range(NR):
                    if not (j >> r) & 1:
                        a, b = phys[j], phys[j | (1 << r)]
                        e(f"        {{ const float2 p0 = vp[{a}], p1 = vp[{b}], l0 = vl[{a}], l1 = vl[{b}];")
                        e("          h[0] = fmaf(p0.y, l0.x, fmaf(-p0.x, l0.y, h[0]));")
                        e("          h[1] = fmaf(p1.y, l1.x, fmaf(-p1.x, l1.y, h[1]));")
                        e("          h[2] = fmaf(p0.y, l1.x, fmaf(-p0.x, l1.y, fmaf(p1.y, l0.x, fmaf(-p1.x, l0.y, h[2]))));")
                        e("          h[3] = fmaf(p0.x, l1.x, fmaf(p0.y, l1.y, fmaf(-p1.x, l0.x, fmaf(-p1.y, l0.y, h[3])))); }")
            else:
                e("        float m[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};")
                for j in range(NR):
                    if not (j >> r) & 1:
                        a, b = phys[j], phys[j | (1 << r)]
                        e(f"        {{ const float2 p0 = vp[{a}], p1 = vp[{b}], l0 = vl[{a}], l1 = vl[{b}];")
                        e("          m[0] += p0.x * l0.x + p0.y * l0.y; m[1] += p0.y * l0.x - p0.x * l0.y;")
                        e("          m[2] += p0.x * l1.x + p0.y * l1.y; m[3] += p0.y * l1.x - p0.x * l1.y;")
                        e("          m[4] += p1.x * l0.x + p1.y * l0.y; m[5] += p1.y * l0.x - p1.x * l0.y;")
                        e("          m[6] += p1.x * l1.x + p1.y * l1.y; m[7] += p1.y * l1.x - p1.x * l1.y; }")
            span = "A.B" if ps else "1"
            var = "csample" if ps else "0"
            trace = "herm_trace" if hermitian else "im_trace"
            acc = "h" if hermitian else "m"
            live = [(k, row) for k, row in enumerate(rows) if row >= 0]
            if wlog == 6 and rows4:
                # the rows' per-lane traces summed over the wave four at a time (wave_sum4); lanes
                # 0..3 then hold them and store them with one instruction
                for c0 in range(0, len(live), 4):
                    grp = live[c0:c0 + 4]
                    ts = [f"{trace}(qt + 8 * ({q} + {k} * {span} + {var}), {acc})" for k, _row in grp]
                    ts += ["0.f"] * (4 - len(grp))
                    rws = [row for _k, row in grp] + [-1] * (4 - len(grp))
                    e(f"        {{ const float g4 = wave_sum4({', '.join(ts)});")
                    e("          const int ln = tid & 63;")
                    e("          if (ln < 4) {")
                    e("            const int v = 2 * (ln & 1) + ((ln >> 1) & 1);")
                    e(f"            const int row = v == 0 ? {rws[0]} : v == 1 ? {rws[1]} : v == 2 ? {rws[2]} : {rws[3]};")
                    e("            if (row >= 0) pinf(A.gpart)[(i64)row * A.NS + uslot] = g4;")
                    e("          }")
                    e("        }")
            else:
                for k, row in live:
                    _emit_put(e, f"{trace}(qt + 8 * ({q} + {k} * {span} + {var}), {acc})", row, wlog, "        ")
            if not probe:
                off = f"8 * ({u} + csample)" if ps else f"{8 * u}"
                if ps and multi:
                    e(f"        const int *uu = ut + {off};")
                    for j in range(NR):
                        if not (j >> r) & 1:
                            a, b = phys[j], phys[j | (1 << r)]
                            e(f"        valu_pair_dag(vp[{a}], vp[{b}], uu); valu_pair_dag(vl[{a}], vl[{b}], uu);")
                elif (si_, r) in cq:
                    nm = cq[(si_, r)]
                    e(f"        const float cf[4] = {{{nm}.x, {nm}.y, {nm}.z, {nm}.w}};")
                    prs = [(phys[j], phys[j | (1 << r)]) for j in range(NR) if not (j >> r) & 1]
                    if interleave == 4:
                        # the psi and lambda pairs of two amplitude pairs: four chains interleaved
                        for k in range(0, len(prs) - 1, 2):
                            (a, b), (c, d) = prs[k], prs[k + 1]
                            e(f"        mfma_pair4(vp[{a}], vp[{b}], vl[{a}], vl[{b}], vp[{c}], vp[{d}], "
                              f"vl[{c}], vl[{d}], cf);")
                        if len(prs) % 2:
                            a, b = prs[-1]
                            e(f"        mfma_pair2(vp[{a}], vp[{b}], vl[{a}], vl[{b}], cf);")
                    else:
                        for a, b in prs:
                            if interleave:
                                e(f"        mfma_pair2(vp[{a}], vp[{b}], vl[{a}], vl[{b}], cf);")
                            else:
                                e(f"        mfma_pair(vp[{a}], vp[{b}], cf); mfma_pair(vl[{a}], vl[{b}], cf);")
                else:
                    e(f"        float cf[4]; coeffs_dag(ut + {off}, cf);")
                    for j in range(NR):
                        if not (j >> r) & 1:
                            a, b = phys[j], phys[j | (1 << r)]
                            e(f"        mfma_pair(vp[{a}], vp[{b}], cf); mfma_pair(vl[{a}], vl[{b}], cf);")
            e("      }")
    if direct_store:
        roff, r0 = _register_offsets(sub, R)
        for j in range(NR):
            if r0 is not None:
                if (j >> r0) & 1:
                    continue
                a, b = phys[j], phys[j | (1 << r0)]
                e(f"      {{ const i64 q = (gbase | {roff[j]}LL) >> 1;")
                e(f"        reinterpret_cast<float4 *>(A.psi)[q] = make_float4(vp[{a}].x, vp[{a}].y, vp[{b}].x, vp[{b}].y);")
                e(f"        reinterpret_cast<float4 *>(A.lam)[q] = make_float4(vl[{a}].x, vl[{a}].y, vl[{b}].x, vl[{b}].y); }}")
            else:
                e(f"      reinterpret_cast<float2 *>(A.psi)[gbase | {roff[j]}LL] = vp[{phys[j]}]; "
                  f"reinterpret_cast<float2 *>(A.lam)[gbase | {roff[j]}LL] = vl[{phys[j]}];")
        e("    }")
        return
    for j in range(NR):
        e(f"      lds_at(tp, lb8 ^ {8 * swj[j]}) = vp[{phys[j]}]; lds_at(tl, lb8 ^ {8 * swj[j]}) = vl[{phys[j]}];")
    if wave_sync:  # the next step keeps the wave-index bits (wave_local)
        e("      __builtin_amdgcn_fence(__ATOMIC_RELEASE, \"wavefront\"); __builtin_amdgcn_wave_barrier();")
        e("      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, \"wavefront\");")
    else:
        e("      __syncthreads();")
    e("    }")


# ---------------------------------------------------------------------------
# compile / cache / launch
# ---------------------------------------------------------------------------

_lock = threading.Lock()
_kernels: dict = {}          # program source hash -> (module handle, {kernel name: function handle})
CACHE_DIR = os.environ.get("QANDLE_AMD_JIT_CACHE",
                           os.path.join(os.path.expanduser("~"), ".cache", "qandle_amd", "jit"))
KERNEL = "qs_pass"


def _compile(src: str, name: str) -> bytes:
    lib = L.load()
    opts = (ctypes.c_char_p * len(OPTIONS))(*[o.encode() for o in OPTIONS])
    h = ctypes.c_void_p()
    rc = lib.qs_jit_compile(src.encode(), (name + ".hip").encode(), opts, len(OPTIONS), ctypes.byref(h))
    try:
        if rc != 0:
            ls = lib.qs_jit_log_size(h) if h.value else 0
            log = ""
            if ls and ls > 0:
                buf = ctypes.create_string_buffer(int(ls) + 1)
                lib.qs_jit_log_copy(h, buf, int(ls) + 1)
                log = buf.value.decode(errors="replace")
            raise L.EngineError(f"hipRTC compilation of {name} failed:\n{log[-4000:]}")
        size = lib.qs_jit_code_size(h)
        buf = ctypes.create_string_buffer(int(size))
        L.check(lib.qs_jit_code_copy(h, buf, int(size)), "qs_jit_code_copy")
        return buf.raw
    finally:
        if h.value:
            lib.qs_jit_release_code(h)


def _key(src: str) -> str:
    return hashlib.sha256((PRELUDE + src + " ".join(OPTIONS)).encode()).hexdigest()[:32]


def _code_for(src: str) -> bytes:
    """Code object of one kernel source: from the disk cache, else compiled and stored."""
    key = _key(src)
    path = os.path.join(CACHE_DIR, key + ".hsaco")
    try:
        with open(path, "rb") as f:
            code = f.read()
        if code:
            return code
    except OSError:
        pass
    code = _compile(PRELUDE + src, "qs_jit_" + key)
    try:
        os.makedirs(CACHE_DIR, exist_ok=True)
        tmp = path + f".{os.getpid()}.{threading.get_ident()}.tmp"
        with open(tmp, "wb") as f:
            f.write(code)
        os.replace(tmp, path)
    except OSError:
        pass
    return code


def kernels_for(sources, names):
    """Function handles of the kernels ``names`` defined by ``sources``.

    All sources form one program (hipRTC serialises compilations, and one
    program pays the fixed compiler start-up once); its code object is cached
    on disk by source hash and loaded once per process."""
    src = "\n\n".join(sources)
    key = _key(src)
    with _lock:
        hit = _kernels.get(key)
    if hit is None:
        code = _code_for(src)
        lib = L.load()
        with _lock:
            hit = _kernels.get(key)
            if hit is None:
                mod = ctypes.c_void_p()
                L.check(lib.qs_jit_load(code, len(code), ctypes.byref(mod)), "qs_jit_load")
                fns = {}
                for nm in names:
                    fn = ctypes.c_void_p()
                    L.check(lib.qs_jit_function(mod, nm.encode(), ctypes.byref(fn)), f"qs_jit_function({nm})")
                    fns[nm] = fn
                hit = (mod, fns)
                _kernels[key] = hit
    return [hit[1][nm] for nm in names]


def lane_mode(lanes, pipelined: bool):
    """How a forward program's kernels take their matrix-core operands: "lds"
    (a compact per-pass block, lane_index), True (the whole table, indexed by
    utab entry) or False (scalar reads); pipelined kernels take none."""
    if pipelined or not lanes:
        return False
    return "lds" if lanes == "lds" else True


# A pass that applies at least this many fused unitaries loads its first sub-pass's
# registers straight from the state instead of staging the tile through LDS: such a pass
# is bound by its arithmetic, and the staging's LDS traffic and barrier sit on that path
# (20 qubits x 64: 437 vs 463 us for a 38-unitary pass, 707 vs 755 us for a 68-unitary one,
# results bit for bit the same), while streaming-bound passes lose by it (257 vs 197 us for
# 12 unitaries: its 8-byte loads keep fewer bytes in flight than the 16-byte staging loads).
DIRECT_LOAD_MIN_UNITARIES = 32
# "auto": the last sub-pass stores its registers straight to the state in the generating
# first pass (5-6% faster: the tile never goes back through LDS), the measuring last pass
# (4-8%: it skips the LDS write-back) and any pass whose direct stores fill whole 128-byte
# lines (3-5%); elsewhere a wave's direct stores scatter over up to 64 lines and cost up to 33%
# (24 qubits), so those passes store through the tile.  Results are bit for bit the same.
DIRECT_STORE = "auto"


class ProgramKernels:
    """The specialised kernels of every pass of one encoded forward program.

    ``waves`` = 2 asks for registers that let two waves share a SIMD: the LDS
    of a 2^13-amplitude tile already limits a compute unit to two workgroups,
    and without the bound the compiler spends accumulator registers until a
    measuring pass drops to one (351 vs 256 us for that pass at 20 qubits).
    ``lanes`` = "lds" and ``split_pairs`` (pass_source) are the defaults: 3.48
    vs 3.67 ms per 20-qubit forward, 4.69 vs 4.89 at 24 qubits, results bit
    for bit the same.  ``interleave`` (pass_source) = 4 is a default too:
    within 1% of 2 and of none on these shapes, and each step halves the
    MFMA wait states.  ``async_loads`` (pass_source_async: two tile buffers,
    one workgroup per compute unit) and ``tail_prefetch`` (pass_source_tail:
    the next tile streams in during the last sub-pass) are options that give
    the same bits and measured slower or no faster."""

    def __init__(self, passes, n: int, R: int, TB: int, gen_first: bool, measure: bool,
                 prefetch: bool = False, waves: int = 2, readout=None, valu_pairs: int = 0,
                 grid: str = "full", pipeline: int = 0, lanes="lds", split_pairs: bool = True,
                 interleave: int = 4, wave_sync: bool = True, stagger: int = 0, load: bool = True,
                 async_loads: bool = False, tail_prefetch: bool = False, direct_store=DIRECT_STORE,
                 direct_load=DIRECT_LOAD_MIN_UNITARIES):
        self.n, self.R, self.TB = n, R, TB
        self.pipelined = bool(pipeline) and n >= R + TB
        self.block = (1 << TB) + ((1 << pipeline) if self.pipelined else 0)
        self.full_grid = grid == "full" and not self.pipelined
        npass = len(passes)
        self.sources, self.names = [], []
        self.lane_index = []  # utab index of every lane-table entry (lanes == "lds"), pass by pass
        self.persistent = []  # per pass: True when its kernel walks tiles with a resident grid
        self.direct = []  # per pass: (registers loaded from the state, registers stored to it)
        for i, (w, _nsub, lpos) in enumerate(passes):
            words = w.cpu().numpy() if hasattr(w, "cpu") else w
            nm = f"{KERNEL}_{i}"
            lane_off = len(self.lane_index)
            if lanes == "lds" and not self.pipelined:
                self.lane_index += lane_entries(words, R, TB, n < R + TB)
            # direct_load: True, False, or the fewest fused unitaries a pass needs to load directly;
            # direct_store: True, False or "auto" (DIRECT_STORE_AUTO)
            subs_i = PL.decode_pass(words, R, TB)
            units = sum(len(bits) for sub in subs_i for bits, _cn in sub.segments)
            dload = direct_load if isinstance(direct_load, bool) else (0 < direct_load <= units)
            dstore = direct_store
            if direct_store == "auto":
                dstore = (gen_first and i == 0) or (measure and i == npass - 1) or _full_lines(subs_i[-1], R)
            use_async = async_loads and not self.pipelined and n >= R + TB and not (gen_first and i == 0)
            use_tail = tail_prefetch and not use_async and not self.pipelined and n >= R + TB
            self.persistent.append(use_async or use_tail)
            if use_tail:
                self.sources.append(pass_source_tail(nm, words, n, R, TB, lpos, gen_first and i == 0,
                                                     measure and i == npass - 1, waves, readout, lanes, lane_off,
                                                     interleave, wave_sync))
            elif use_async:
                self.sources.append(pass_source_async(nm, words, n, R, TB, lpos, measure and i == npass - 1, waves,
                                                      readout, lanes, lane_off, interleave, wave_sync))
            elif self.pipelined:
                self.sources.append(pass_source_pipelined(nm, words, n, R, TB, lpos, gen_first and i == 0,
                                                          measure and i == npass - 1, waves, readout, pipeline))
            else:
                self.sources.append(pass_source(nm, words, n, R, TB, lpos, gen_first and i == 0,
                                                measure and i == npass - 1, prefetch, waves, readout, valu_pairs,
                                                lanes, split_pairs, lane_off, interleave, wave_sync, stagger,
                                                direct_store=dstore, direct_load=dload))
            plain, one = not (use_tail or use_async or self.pipelined), n >= R + TB  # pass_source's gating
            self.direct.append((plain and one and bool(dload) and not (gen_first and i == 0) and not prefetch,
                                plain and one and bool(dstore)))
            self.names.append(nm)
        if load:
            self.funcs = kernels_for(self.sources, self.names)
        else:  # compile into the code-object cache only (a process that launches nothing)
            _code_for("\n\n".join(self.sources))
            self.funcs = []
        self.grid_cap = [None] * len(self.funcs)
        self.lanes = lane_mode(lanes, self.pipelined)
        self._lane_idx = {}

    def lane_index_on(self, device):
        """The lane-table index list as a device int32 tensor (None unless lanes == "lds")."""
        if self.lanes != "lds" or not self.lane_index:
            return None
        key = str(device)
        t = self._lane_idx.get(key)
        if t is None:
            import torch

            t = L.upload(self.lane_index, torch.int32, device)
            self._lane_idx[key] = t
        return t

    def launch(self, i: int, state, src, utab, cs, efac, mpart, B: int, ntiles: int, store: int, stream,
               ltab=None):
        lib = L.load()
        if self.full_grid and not self.persistent[i]:
            grid = max(1, min(ntiles, 0x7fffffff))
        else:
            if self.grid_cap[i] is None:
                g = lib.qs_jit_resident_blocks(self.funcs[i], self.block)
                self.grid_cap[i] = int(g) if g and g > 0 else 2048
            grid = max(1, min(ntiles, self.grid_cap[i]))
        args = struct.pack(ARGS_FMT, state or 0, src or 0, utab or 0, cs or 0, efac or 0, mpart or 0, ltab or 0,
                           B, ntiles, store, 0)
        buf = ctypes.create_string_buffer(args, len(args))
        L.check(lib.qs_jit_launch(self.funcs[i], grid, self.block, buf, len(args), stream), "qs_jit_launch")


class AdjointKernels:
    """The specialised kernels of every step of one encoded adjoint sweep.

    ``direct_load`` / ``direct_store`` follow the forward's per-pass rules (a
    step of 32+ fused unitaries loads psi and lambda straight from the state; a
    step whose last sub-pass's stores fill whole 128-byte lines stores them
    straight back): 4.07-4.08 vs 4.17-4.20 ms for the 15 steps at 16 qubits x
    256 x 20 layers, gradients bit for bit the same.  ``interleave`` = 4 (two
    amplitude pairs' psi and lambda chains at a time) gives the same bits and
    measured 1.5% slower than the default two chains (4.21-4.24 vs 4.15-4.16 ms).

    Defaults as for ProgramKernels (two waves per SIMD, the U^dagger lane
    table in LDS, 8-byte pair moves): 18.1 vs 22.2 ms per forward + backward
    at 16 qubits x 256 x 20 layers, 27.5 vs 29.1 ms at 20 qubits x 64,
    gradients bit for bit the same."""

    def __init__(self, passes, n: int, R: int, TB: int, seed_first: bool, prefetch: bool = False,
                 waves: int = 2, readout=None, grid: str = "full", lanes="lds", split_pairs: bool = True,
                 interleave: bool = True, wave_sync: bool = True, hermitian: bool = True,
                 direct_load=DIRECT_LOAD_MIN_UNITARIES, direct_store=DIRECT_STORE, rows4: bool = True):
        self.block = 1 << TB
        self.full_grid = grid == "full"
        self.sources, self.names = [], []
        self.lane_index = []
        for i, (w, _nsub, lpos) in enumerate(passes):
            words = w.cpu().numpy() if hasattr(w, "cpu") else w
            nm = f"qs_adj_{i}"
            lane_off = len(self.lane_index)
            if lanes == "lds":
                self.lane_index += adjoint_lane_entries(words, R, TB)
            # direct_load: True, False or the fewest fused unitaries a step needs to load directly;
            # direct_store: True, False or "auto" (whole 128-byte lines, as for the forward)
            subs_i = PL.decode_adjoint_pass(words, R, TB)
            units = sum(len(bits) for sub in subs_i for _cn, bits in sub.segments)
            dload = direct_load if isinstance(direct_load, bool) else (0 < direct_load <= units)
            dstore = _full_lines(subs_i[-1], R) if direct_store == "auto" else direct_store
            self.sources.append(adjoint_source(nm, words, n, R, TB, lpos, seed_first and i == 0, prefetch, waves,
                                               readout, lanes, split_pairs, lane_off, interleave, wave_sync,
                                               hermitian, direct_load=dload, direct_store=dstore, rows4=rows4))
            self.names.append(nm)
        self.funcs = kernels_for(self.sources, self.names)
        self.grid_cap = [None] * len(self.funcs)

    def lane_index_on(self, device):
        """The U^dagger lane-table index list as a device int32 tensor (None when unused)."""
        if not self.lane_index:
            return None
        key = str(device)
        cache = self.__dict__.setdefault("_lane_idx", {})
        if key not in cache:
            import torch

            cache[key] = L.upload(self.lane_index, torch.int32, device)
        return cache[key]

    def launch(self, i: int, psi, lam, utab, qtab, cs, gpart, seed_w, B: int, NS: int, ntiles: int, stream,
               ltab=None, psi_src=None):
        """``psi_src``: where the step reads psi (None: psi itself); it writes psi."""
        lib = L.load()
        if self.full_grid:
            grid = max(1, min(ntiles, 0x7fffffff))
        else:
            if self.grid_cap[i] is None:
                g = lib.qs_jit_resident_blocks(self.funcs[i], self.block)
                self.grid_cap[i] = int(g) if g and g > 0 else 2048
            grid = max(1, min(ntiles, self.grid_cap[i]))
        args = struct.pack(ADJ_FMT, psi, lam, utab or 0, qtab or 0, cs or 0, gpart, seed_w or 0, ltab or 0,
                           psi_src or psi, B, NS, ntiles)
        buf = ctypes.create_string_buffer(args, len(args))
        L.check(lib.qs_jit_launch(self.funcs[i], grid, self.block, buf, len(args), stream), "qs_jit_launch")
