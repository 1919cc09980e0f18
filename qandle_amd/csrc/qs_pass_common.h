// Pieces shared by the fused forward pass kernel and the adjoint pass kernel:
// the sub-pass header layout, the LDS swizzle and the compile-time-indexed
// register operations behind the uniform op dispatch.
#pragma once

#include "qs_common.h"
#include "qs_program.h"

namespace qs {

// sub-pass header layout (words); see qs_program.h for the op words
constexpr int kHdr = 56;
constexpr int kHdrP = 1;      // [1, 1+R): register local bits
constexpr int kHdrSwz = 8;    // [8, 8+R): swizzled LDS stride of each register bit
constexpr int kHdrGP = 16;    // [16, 16+R): global position of each register bit
constexpr int kHdrG = 24;     // [24, 24+TB): thread-bit local bits
constexpr int kHdrGG = 40;    // [40, 40+TB): global position of each thread bit


// XOR swizzle of an LDS amplitude index: folds the bits >= 5 into bits 0..4
// so that lanes that differ only in high local bits land in different banks.
// Linear over XOR (so register addresses are base ^ stride chains) and a
// bijection (bits >= 5 are unchanged).  An amplitude pair (2m, 2m+1) stays in
// one 16-byte slot, in swapped order when bit 0 of swz(2m) is set.
__host__ __device__ __forceinline__ int swz(int l) {
  int f = (l >> 5) ^ (l >> 10) ^ (l >> 15);
  return l ^ (f & 31);
}

// Address-space views of the buffers a descriptor points to.  A pointer read
// from memory (a work item's descriptor) is generic: its loads and stores are
// flat instructions, which count against the LDS counter too, so every LDS
// wait would also wait for the next tile's HBM loads.  The state buffers are
// global memory; the programs and tables are read-only while a kernel runs,
// and reading them through the constant address space puts uniform reads on
// the scalar path (and the branches on their words on scalar branches).
template <class T>
using gptr = __attribute__((address_space(1))) T *;
template <class T>
using kptr = const __attribute__((address_space(4))) T *;
template <class T>
__device__ __forceinline__ gptr<T> to_global(T *p) {
  return (gptr<T>)p;
}
template <class T>
__device__ __forceinline__ kptr<T> to_const(const T *p) {
  return (kptr<T>)p;
}

// ---------------------------------------------------------------------------
// Per-bit operations.  A 2x2 complex matrix is 8 floats
// (u00.re, u00.im, u01.re, u01.im, u10.re, u10.im, u11.re, u11.im), column
// convention: (a0, a1) <- (u00 a0 + u01 a1, u10 a0 + u11 a1).
//
// The pass kernels hold their read-only tables (fused unitaries, gradient
// generators, phase coefficients) as int32 pointers and convert the bits with
// __int_as_float: no float store of a kernel can then alias a table read, so
// the compiler keeps these uniform reads on the scalar (SMEM) path instead of
// issuing vector loads that wait behind the tile's HBM traffic.
// ---------------------------------------------------------------------------
struct M8 {
  float x[8];
};

template <class P>
__device__ __forceinline__ M8 load_m8(P p) {
  M8 m;
#pragma unroll
  for (int i = 0; i < 8; ++i) m.x[i] = __int_as_float(p[i]);
  return m;
}

// A fused unitary on the vector ALU (two complex multiply-adds per output);
// used where the four lanes of a matrix-core block may hold different
// unitaries (per-sample unitaries in a tile of several samples).
template <int BIT, int NR>
__device__ __forceinline__ void apply_valu(float2 (&v)[NR], const M8 &u) {
  const float a = u.x[0], b = u.x[1], c = u.x[2], d = u.x[3], e = u.x[4], f = u.x[5], g = u.x[6], h = u.x[7];
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    if (!(j & (1 << BIT))) {
      const float2 x = v[j], y = v[j | (1 << BIT)];
      v[j] = make_float2(a * x.x - b * x.y + c * y.x - d * y.y, a * x.y + b * x.x + c * y.y + d * y.x);
      v[j | (1 << BIT)] = make_float2(e * x.x - f * x.y + g * y.x - h * y.y, e * x.y + f * x.x + g * y.y + h * y.x);
    }
  }
}

__device__ __forceinline__ M8 dagger(const M8 &u) {
  M8 t;
  t.x[0] = u.x[0]; t.x[1] = -u.x[1]; t.x[2] = u.x[4]; t.x[3] = -u.x[5];
  t.x[4] = u.x[2]; t.x[5] = -u.x[3]; t.x[6] = u.x[6]; t.x[7] = -u.x[7];
  return t;
}

// A fused unitary runs on the matrix cores.  In the real form (re0, im0, re1,
// im1) of an amplitude pair, U is the 4x4 matrix
//   [[a, -b, c, -d], [b, a, d, c], [e, -f, g, -h], [f, e, h, g]]
// (u00 = a + ib, u01 = c + id, u10 = e + if, u11 = g + ih) and U^dagger is its
// transpose.  v_mfma_f32_4x4x1f32 computes, in 16 blocks of 4 lanes,
// D[i][j] += A(lane 4b + i) * B(lane 4b + j) with output row i in accumulator
// register i of lane 4b + j: with A_k(lane) = U[lane & 3][k] and B_k(lane) =
// component k of the lane's own pair, four steps k = 0..3 give every lane
// U times its pair.  The matrix units run at twice the FP32 rate of the
// vector ALU and leave it free for the rest of the pass.
typedef float qs_f4 __attribute__((ext_vector_type(4)));

// 16-byte state moves and 8-byte table reads through the address-space views
// (the HIP vector classes have no operators for address-space-qualified
// objects; clang's native vectors and scalars do)
__device__ __forceinline__ float4 load_g4(const float *p, int64_t i) {
  const qs_f4 v = to_global(reinterpret_cast<const qs_f4 *>(p))[i];
  return make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void store_g4(float *p, int64_t i, const float4 &v) {
  qs_f4 x = {v.x, v.y, v.z, v.w};
  to_global(reinterpret_cast<qs_f4 *>(p))[i] = x;
}
__device__ __forceinline__ float2 load_k2(kptr<int32_t> p, int64_t i) {
  return make_float2(__int_as_float(p[2 * i]), __int_as_float(p[2 * i + 1]));
}

struct MU {
  float a[4];  // A_k of this lane, k = 0..3
};

// Coefficients of U (or of U^dagger when DAG) for the lane's matrix row.
template <bool DAG>
__device__ __forceinline__ MU mfma_coeffs(const M8 &u) {
  const float a = u.x[0], b = u.x[1], c = u.x[2], d = u.x[3], e = u.x[4], f = u.x[5], g = u.x[6], h = u.x[7];
  const int row = threadIdx.x & 3;
  const bool r1 = row & 1, r2 = row & 2;
  MU m;
  if (!DAG) {  // A_k = U[row][k]
    m.a[0] = r2 ? (r1 ? f : e) : (r1 ? b : a);
    m.a[1] = r2 ? (r1 ? e : -f) : (r1 ? a : -b);
    m.a[2] = r2 ? (r1 ? h : g) : (r1 ? d : c);
    m.a[3] = r2 ? (r1 ? g : -h) : (r1 ? c : -d);
  } else {  // A_k = U[k][row]
    m.a[0] = r2 ? (r1 ? -d : c) : (r1 ? -b : a);
    m.a[1] = r2 ? (r1 ? c : d) : (r1 ? a : b);
    m.a[2] = r2 ? (r1 ? -h : g) : (r1 ? -f : e);
    m.a[3] = r2 ? (r1 ? g : h) : (r1 ? e : f);
  }
  return m;
}

template <int BIT, int NR>
__device__ __forceinline__ void apply_mfma(float2 (&v)[NR], const MU &m) {
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    if (!(j & (1 << BIT))) {
      const float2 x = v[j], y = v[j | (1 << BIT)];
      qs_f4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_4x4x1f32(m.a[0], x.x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_4x4x1f32(m.a[1], x.y, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_4x4x1f32(m.a[2], y.x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_4x4x1f32(m.a[3], y.y, acc, 0, 0, 0);
      v[j] = make_float2(acc[0], acc[1]);
      v[j | (1 << BIT)] = make_float2(acc[2], acc[3]);
    }
  }
}

// The cross density M_ab = sum psi_a conj(lam_b) of bit BIT over the thread's
// registers enters a gradient only through Im Tr(Q M), and every gradient
// generator Q (a Pauli transported through the rest of its fused unitary,
// V P V^dagger) is Hermitian: Q00 and Q11 are real and Q10 = conj(Q01), so
// Im Tr(Q M) = Q00 Im M00 + Q11 Im M11 + Re Q01 (Im M01 + Im M10)
//            + Im Q10 (Re M01 - Re M10).
// herm_density accumulates those four combinations (four fused multiply-add
// chains instead of the eight entries of M).
struct H4 {
  float h[4];
};

template <int BIT, int NR>
__device__ __forceinline__ H4 herm_density(const float2 (&p)[NR], const float2 (&l)[NR]) {
  float h0 = 0.f, h1 = 0.f, h2 = 0.f, h3 = 0.f;
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    if (!(j & (1 << BIT))) {
      const float2 p0 = p[j], p1 = p[j | (1 << BIT)], l0 = l[j], l1 = l[j | (1 << BIT)];
      h0 = fmaf(p0.y, l0.x, fmaf(-p0.x, l0.y, h0));
      h1 = fmaf(p1.y, l1.x, fmaf(-p1.x, l1.y, h1));
      h2 = fmaf(p0.y, l1.x, fmaf(-p0.x, l1.y, fmaf(p1.y, l0.x, fmaf(-p1.x, l0.y, h2))));
      h3 = fmaf(p0.x, l1.x, fmaf(p0.y, l1.y, fmaf(-p1.x, l0.x, fmaf(-p1.y, l0.y, h3))));
    }
  }
  return H4{{h0, h1, h2, h3}};
}

// Im Tr(Q M) from herm_density for a Hermitian generator stored as 8 floats
// (q00.re, q00.im, q01.re, q01.im, q10.re, q10.im, q11.re, q11.im), int bits.
template <class P>
__device__ __forceinline__ float herm_trace(P q, const H4 &m) {
  return (__int_as_float(q[0]) * m.h[0] + __int_as_float(q[6]) * m.h[1]) +
         (__int_as_float(q[4]) * m.h[2] + __int_as_float(q[5]) * m.h[3]);
}

// Wave reductions without LDS traffic.  dpp_add<CTRL> adds the value of the
// lane a DPP control selects: quad_perm [1,0,3,2] (0xB1) and [2,3,0,1] (0x4E)
// exchange within quads, row_ror 4 (0x124) and 8 (0x128) rotate a 16-lane
// row.  The gfx950 row and half swaps (permlane16_swap, permlane32_swap) with
// the same value as both operands return the value's two halves, whose sum
// is s[l] + s[l ^ 16] (s[l] + s[l ^ 32]) in every lane.
template <int CTRL>
__device__ __forceinline__ float dpp_add(float s) {
  return s + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(s), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float swap16_add(float s) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(s), __float_as_uint(s), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float swap32_add(float s) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(s), __float_as_uint(s), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// Four values summed over a 64-lane wave together: two exchange rounds (lane
// l ^ 1, then l ^ 2) leave each lane one value's quad sum, value index
// 2 (l & 1) + ((l >> 1) & 1); the row rotations add a row's lanes of the same
// index and the two swaps the four rows.  Lanes 0, 2, 1, 3 (and every lane of
// the same index) end with the wave sums of values 0, 1, 2, 3.
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
// the value index wave_sum4 leaves in lane l
__device__ __forceinline__ int wave_sum4_index(int l) { return 2 * (l & 1) + ((l >> 1) & 1); }

// run-time bit -> compile-time template (the switch folds away once the
// caller's loop over register bits is unrolled)
template <int R>
__device__ __forceinline__ void apply_valu_at(float2 (&v)[1 << R], int r, const M8 &u) {
  switch (r) {
    case 0: apply_valu<0>(v, u); break;
    case 1: if constexpr (R > 1) apply_valu<1>(v, u); break;
    case 2: if constexpr (R > 2) apply_valu<2>(v, u); break;
    case 3: if constexpr (R > 3) apply_valu<3>(v, u); break;
    case 4: if constexpr (R > 4) apply_valu<4>(v, u); break;
    default: break;
  }
}
template <int R>
__device__ __forceinline__ void apply_mfma_at(float2 (&v)[1 << R], int r, const MU &m) {
  switch (r) {
    case 0: apply_mfma<0>(v, m); break;
    case 1: if constexpr (R > 1) apply_mfma<1>(v, m); break;
    case 2: if constexpr (R > 2) apply_mfma<2>(v, m); break;
    case 3: if constexpr (R > 3) apply_mfma<3>(v, m); break;
    case 4: if constexpr (R > 4) apply_mfma<4>(v, m); break;
    default: break;
  }
}
template <int R>
__device__ __forceinline__ H4 herm_density_at(const float2 (&p)[1 << R], const float2 (&l)[1 << R], int r) {
  switch (r) {
    case 1: if constexpr (R > 1) return herm_density<1>(p, l); break;
    case 2: if constexpr (R > 2) return herm_density<2>(p, l); break;
    case 3: if constexpr (R > 3) return herm_density<3>(p, l); break;
    case 4: if constexpr (R > 4) return herm_density<4>(p, l); break;
    default: break;
  }
  return herm_density<0>(p, l);
}

// Per-thread phase of a sub-pass's RZ gates on non-register bits:
// prod_k e^{-/+ i phi_k/2} by the thread's value of bit pos_k.  cs holds the
// (cos, sin) float pairs as int32 bits.
template <class PH, class CS>
__device__ __forceinline__ float2 phase_product(PH ph, int nphase, CS cs, int64_t csample, int64_t gbase,
                                                bool inverse) {
  float pr = 1.f, pi = 0.f;
  for (int k = 0; k < nphase; ++k) {
    const int fl = ph[4 * k + 2];
    if (fl & QS_SEG_NOUNAPPLY) continue;
    const int64_t idx = ph[4 * k + 1] + ((fl & QS_SEG_PER_SAMPLE) ? csample : 0);
    const float c = __int_as_float(cs[2 * idx]), s = __int_as_float(cs[2 * idx + 1]);
    float sn = ((gbase >> ph[4 * k]) & 1) ? -s : s;
    if (inverse) sn = -sn;
    const float nr = pr * c + pi * sn, ni = pi * c - pr * sn;  // (pr + i pi)(c - i sn)
    pr = nr;
    pi = ni;
  }
  return make_float2(pr, pi);
}

template <int CB, int TBIT, int NR>
__device__ __forceinline__ void cnot_reg(float2 (&v)[NR]) {
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    if ((j & (1 << CB)) && !(j & (1 << TBIT))) {
      float2 t = v[j];
      v[j] = v[j | (1 << TBIT)];
      v[j | (1 << TBIT)] = t;
    }
  }
}

template <int TBIT, int NR>
__device__ __forceinline__ void swap_reg(float2 (&v)[NR]) {
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    if (!(j & (1 << TBIT))) {
      float2 t = v[j];
      v[j] = v[j | (1 << TBIT)];
      v[j | (1 << TBIT)] = t;
    }
  }
}

template <int R>
__device__ __forceinline__ void cnot_dispatch(float2 (&v)[1 << R], int cb, int tb) {
#define QS_CN(C, T)                                              \
  case (C) * 8 + (T):                                            \
    if constexpr ((C) < R && (T) < R && (C) != (T)) cnot_reg<C, T>(v); \
    break;
  switch (cb * 8 + tb) {
    QS_CN(0, 1) QS_CN(0, 2) QS_CN(0, 3) QS_CN(0, 4)
    QS_CN(1, 0) QS_CN(1, 2) QS_CN(1, 3) QS_CN(1, 4)
    QS_CN(2, 0) QS_CN(2, 1) QS_CN(2, 3) QS_CN(2, 4)
    QS_CN(3, 0) QS_CN(3, 1) QS_CN(3, 2) QS_CN(3, 4)
    QS_CN(4, 0) QS_CN(4, 1) QS_CN(4, 2) QS_CN(4, 3)
    default: break;
  }
#undef QS_CN
}

template <int R>
__device__ __forceinline__ void swap_dispatch(float2 (&v)[1 << R], int tb) {
  switch (tb) {
    case 0: swap_reg<0>(v); break;
    case 1: if constexpr (R > 1) swap_reg<1>(v); break;
    case 2: if constexpr (R > 2) swap_reg<2>(v); break;
    case 3: if constexpr (R > 3) swap_reg<3>(v); break;
    case 4: if constexpr (R > 4) swap_reg<4>(v); break;
    default: break;
  }
}

// One CNOT word: bit 0 = control outside the registers (its value is uniform
// per thread), bits 8-15 = target register, bits 16+ = control (register
// index, or global bit position when uniform).
template <int R>
__device__ __forceinline__ void run_cnot(float2 (&v)[1 << R], int cw, int64_t gbase) {
  const int tgt = (cw >> 8) & 0xff, ctl = cw >> 16;
  if (!(cw & 1))
    cnot_dispatch<R>(v, ctl, tgt);
  else if ((gbase >> ctl) & 1)
    swap_dispatch<R>(v, tgt);
}

}  // namespace qs
