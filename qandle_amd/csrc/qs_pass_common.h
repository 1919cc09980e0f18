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

// ---------------------------------------------------------------------------
// v2 per-bit operations.  A 2x2 complex matrix is 8 floats
// (u00.re, u00.im, u01.re, u01.im, u10.re, u10.im, u11.re, u11.im), column
// convention: (a0, a1) <- (u00 a0 + u01 a1, u10 a0 + u11 a1).
// ---------------------------------------------------------------------------
template <int BIT, int NR>
__device__ __forceinline__ void apply_u(float2 (&v)[NR], const float *u) {
  const float a = u[0], b = u[1], c = u[2], d = u[3], e = u[4], f = u[5], g = u[6], h = u[7];
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    if (!(j & (1 << BIT))) {
      const float2 x = v[j], y = v[j | (1 << BIT)];
      v[j] = make_float2(a * x.x - b * x.y + c * y.x - d * y.y, a * x.y + b * x.x + c * y.y + d * y.x);
      v[j | (1 << BIT)] = make_float2(e * x.x - f * x.y + g * y.x - h * y.y, e * x.y + f * x.x + g * y.y + h * y.x);
    }
  }
}

// (a0, a1) <- U^dagger (a0, a1) = (conj(u00) a0 + conj(u10) a1, conj(u01) a0 + conj(u11) a1)
template <int BIT, int NR>
__device__ __forceinline__ void apply_udag(float2 (&v)[NR], const float *u) {
  const float a = u[0], b = -u[1], c = u[4], d = -u[5], e = u[2], f = -u[3], g = u[6], h = -u[7];
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    if (!(j & (1 << BIT))) {
      const float2 x = v[j], y = v[j | (1 << BIT)];
      v[j] = make_float2(a * x.x - b * x.y + c * y.x - d * y.y, a * x.y + b * x.x + c * y.y + d * y.x);
      v[j | (1 << BIT)] = make_float2(e * x.x - f * x.y + g * y.x - h * y.y, e * x.y + f * x.x + g * y.y + h * y.x);
    }
  }
}

// Cross density of bit BIT over the thread's registers: M_ab = sum psi_a conj(lam_b),
// 8 floats (m00.re, m00.im, m01.re, m01.im, m10.re, m10.im, m11.re, m11.im).
template <int BIT, int NR>
__device__ __forceinline__ void cross_density(const float2 (&p)[NR], const float2 (&l)[NR], float *m) {
  float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f, m4 = 0.f, m5 = 0.f, m6 = 0.f, m7 = 0.f;
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    if (!(j & (1 << BIT))) {
      const float2 p0 = p[j], p1 = p[j | (1 << BIT)], l0 = l[j], l1 = l[j | (1 << BIT)];
      // psi_a * conj(lam_b) = (pa.x lb.x + pa.y lb.y) + i (pa.y lb.x - pa.x lb.y)
      m0 += p0.x * l0.x + p0.y * l0.y;  m1 += p0.y * l0.x - p0.x * l0.y;
      m2 += p0.x * l1.x + p0.y * l1.y;  m3 += p0.y * l1.x - p0.x * l1.y;
      m4 += p1.x * l0.x + p1.y * l0.y;  m5 += p1.y * l0.x - p1.x * l0.y;
      m6 += p1.x * l1.x + p1.y * l1.y;  m7 += p1.y * l1.x - p1.x * l1.y;
    }
  }
  m[0] = m0; m[1] = m1; m[2] = m2; m[3] = m3; m[4] = m4; m[5] = m5; m[6] = m6; m[7] = m7;
}

// run-time bit -> compile-time template (the switch folds away once the
// caller's loop over register bits is unrolled)
template <int R>
__device__ __forceinline__ void apply_u_at(float2 (&v)[1 << R], int r, const float *u) {
  switch (r) {
    case 0: apply_u<0>(v, u); break;
    case 1: if constexpr (R > 1) apply_u<1>(v, u); break;
    case 2: if constexpr (R > 2) apply_u<2>(v, u); break;
    case 3: if constexpr (R > 3) apply_u<3>(v, u); break;
    case 4: if constexpr (R > 4) apply_u<4>(v, u); break;
    default: break;
  }
}
template <int R>
__device__ __forceinline__ void apply_udag_at(float2 (&v)[1 << R], int r, const float *u) {
  switch (r) {
    case 0: apply_udag<0>(v, u); break;
    case 1: if constexpr (R > 1) apply_udag<1>(v, u); break;
    case 2: if constexpr (R > 2) apply_udag<2>(v, u); break;
    case 3: if constexpr (R > 3) apply_udag<3>(v, u); break;
    case 4: if constexpr (R > 4) apply_udag<4>(v, u); break;
    default: break;
  }
}
template <int R>
__device__ __forceinline__ void cross_density_at(const float2 (&p)[1 << R], const float2 (&l)[1 << R], int r, float *m) {
  switch (r) {
    case 0: cross_density<0>(p, l, m); break;
    case 1: if constexpr (R > 1) cross_density<1>(p, l, m); break;
    case 2: if constexpr (R > 2) cross_density<2>(p, l, m); break;
    case 3: if constexpr (R > 3) cross_density<3>(p, l, m); break;
    case 4: if constexpr (R > 4) cross_density<4>(p, l, m); break;
    default: break;
  }
}

// Per-thread phase of a sub-pass's RZ gates on non-register bits:
// prod_k e^{-/+ i phi_k/2} by the thread's value of bit pos_k.
__device__ __forceinline__ float2 phase_product(const int32_t *ph, int nphase, const float2 *cs, int64_t csample,
                                                int64_t gbase, bool inverse) {
  float pr = 1.f, pi = 0.f;
  for (int k = 0; k < nphase; ++k) {
    const int fl = ph[4 * k + 2];
    if (fl & QS_SEG_NOUNAPPLY) continue;
    const float2 c_s = cs[ph[4 * k + 1] + ((fl & QS_SEG_PER_SAMPLE) ? csample : 0)];
    float sn = ((gbase >> ph[4 * k]) & 1) ? -c_s.y : c_s.y;
    if (inverse) sn = -sn;
    const float nr = pr * c_s.x + pi * sn, ni = pi * c_s.x - pr * sn;  // (pr + i pi)(c - i sn)
    pr = nr;
    pi = ni;
  }
  return make_float2(pr, pi);
}

// Im Tr(Q M) = Im sum_ab Q_ab M_ba
__device__ __forceinline__ float im_trace(const float *q, const float *m) {
  // Q00 M00 + Q01 M10 + Q10 M01 + Q11 M11, imaginary parts
  return (q[0] * m[1] + q[1] * m[0]) + (q[2] * m[5] + q[3] * m[4]) +
         (q[4] * m[3] + q[5] * m[2]) + (q[6] * m[7] + q[7] * m[6]);
}

template <int KIND, int BIT, int NR>
__device__ __forceinline__ void rot_reg(float2 (&v)[NR], float c, float s) {
#pragma unroll
  for (int j = 0; j < NR; ++j)
    if (!(j & (1 << BIT))) rot_pair<KIND>(c, s, v[j], v[j | (1 << BIT)]);
}

template <int KIND, int R>
__device__ __forceinline__ void rot_dispatch(float2 (&v)[1 << R], int bit, float c, float s) {
  switch (bit) {
    case 0: rot_reg<KIND, 0>(v, c, s); break;
    case 1: if constexpr (R > 1) rot_reg<KIND, 1>(v, c, s); break;
    case 2: if constexpr (R > 2) rot_reg<KIND, 2>(v, c, s); break;
    case 3: if constexpr (R > 3) rot_reg<KIND, 3>(v, c, s); break;
    case 4: if constexpr (R > 4) rot_reg<KIND, 4>(v, c, s); break;
    default: break;
  }
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

// Applies the ops of one sub-pass to the thread's registers.
template <int R>
__device__ __forceinline__ void run_ops(float2 (&v)[1 << R], const int32_t *op, int nops,
                                        const float2 *cs, int64_t sample, int64_t gbase) {
  for (int o = 0; o < nops; ++o, op += 4) {
    const int code = op[0];
    const int a = op[1], b = op[2];
    const int kind = code & 0xff;
    if (kind == QS_OP_CNOT) {
      cnot_dispatch<R>(v, b, a);
      continue;
    }
    if (kind == QS_OP_CNOT_UNIFORM) {
      if ((gbase >> b) & 1) swap_dispatch<R>(v, a);
      continue;
    }
    const int64_t idx = op[3] + ((code & QS_OP_PER_SAMPLE) ? sample : 0);
    const float2 c_s = cs[idx];
    if (kind == QS_OP_RX) {
      rot_dispatch<QS_RX, R>(v, a, c_s.x, c_s.y);
    } else if (kind == QS_OP_RY) {
      rot_dispatch<QS_RY, R>(v, a, c_s.x, c_s.y);
    } else if (kind == QS_OP_RZ) {
      rot_dispatch<QS_RZ, R>(v, a, c_s.x, c_s.y);
    } else {  // QS_OP_RZ_UNIFORM: phase e^{-i phi/2} on bit 0, e^{+i phi/2} on bit 1
      const float c = c_s.x;
      const float s = ((gbase >> a) & 1) ? -c_s.y : c_s.y;
#pragma unroll
      for (int j = 0; j < (1 << R); ++j) {
        float2 x = v[j];
        v[j] = make_float2(c * x.x + s * x.y, c * x.y - s * x.x);
      }
    }
  }
}


}  // namespace qs
