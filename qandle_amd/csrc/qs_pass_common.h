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


// XOR swizzle of an LDS amplitude index: folds bits >= 5 into bits 1..4 so
// that lanes that differ only in high local bits land in different banks.
// Linear over XOR and keeps bit 0 (amplitude pairs stay 16-byte vectors).
__host__ __device__ __forceinline__ int swz(int l) {
  int f = (l >> 5) ^ (l >> 9) ^ (l >> 13);
  return l ^ ((f & 15) << 1);
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
