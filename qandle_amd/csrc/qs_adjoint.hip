// Adjoint (reverse-mode) backward pass kernel for gfx950.
//
// The reference differentiates through torch autograd: every dense
// `state @ M` and every rebuilt 2^n x 2^n matrix is kept on the tape and the
// backward forms grad_M = state^H @ g, an outer product per gate
// (src/qandle/operators.py:272-298, implicit autograd).  Here the backward is
// a reverse sweep over the same fused passes.  A tile holds both the state psi
// and the co-state lambda (the gradient w.r.t. the state) in LDS; for each gate
// in reverse order the kernel
//   1. reduces Im<lambda | P | psi> over the tile (at the gate's output; the
//      gate is exp(-i phi P / 2), so dL/dphi = 1/2 Im<lambda|P|psi>), and
//   2. un-applies the gate to both psi and lambda (U^dagger = U(-phi); CNOT is
//      its own inverse).
// Reductions go per (gate, tile segment) to a workspace in a fixed order;
// qs_adjoint_finish sums them per sample in double.  No float atomics, so
// gradients are bit-reproducible.
#include "qs_pass_common.h"

namespace qs {

struct AdjArgs {
  float *psi;
  float *lam;
  const float *cs;
  float *gpart;
  const int32_t *words;
  int64_t B;
  int64_t NS;  // reduction slots per gradient row
  int n;
  int nsub;
  int nn;
  int32_t lpos[QS_MAX_LOCAL_BITS];
  int32_t npos[QS_MAX_QUBITS];
};

template <int KIND, int BIT, int NR>
__device__ __forceinline__ float im_reg(const float2 (&l)[NR], const float2 (&p)[NR]) {
  float acc = 0.f;
#pragma unroll
  for (int j = 0; j < NR; ++j)
    if (!(j & (1 << BIT))) acc += pauli_im_pair<KIND>(l[j], l[j | (1 << BIT)], p[j], p[j | (1 << BIT)]);
  return acc;
}

template <int KIND, int R>
__device__ __forceinline__ float im_dispatch(const float2 (&l)[1 << R], const float2 (&p)[1 << R], int bit) {
  switch (bit) {
    case 0: return im_reg<KIND, 0>(l, p);
    case 1: if constexpr (R > 1) return im_reg<KIND, 1>(l, p); break;
    case 2: if constexpr (R > 2) return im_reg<KIND, 2>(l, p); break;
    case 3: if constexpr (R > 3) return im_reg<KIND, 3>(l, p); break;
    case 4: if constexpr (R > 4) return im_reg<KIND, 4>(l, p); break;
    default: break;
  }
  return 0.f;
}

template <int R, int TB, bool MULTI>
__global__ void __launch_bounds__(1 << TB) k_adjoint_pass(AdjArgs A) {
  constexpr int T = 1 << TB;
  constexpr int K = R + TB;
  constexpr int NR = 1 << R;
  __shared__ __align__(16) float2 tp[1 << K];
  __shared__ __align__(16) float2 tl[1 << K];

  const int tid = threadIdx.x;
  const int n = A.n;
  const int64_t total = A.B << n;
  const int64_t t = blockIdx.x;
  int64_t base;
  if (MULTI) {
    base = t << K;
  } else {
    const int64_t tlo = t & ((int64_t(1) << A.nn) - 1);
    base = (t >> A.nn) << n;
    for (int i = 0; i < A.nn; ++i)
      if ((tlo >> i) & 1) base |= int64_t(1) << A.npos[i];
  }
  const int64_t block_sample = base >> n;
  const int seglog = MULTI ? (n - R) : TB;
  const int wlog = seglog < 6 ? seglog : 6;
  const int64_t u = t * (T >> wlog) + (tid >> wlog);  // reduction slot of this thread's segment
  const bool leader = (tid & ((1 << wlog) - 1)) == 0;

  int64_t g_tid = 0;
#pragma unroll
  for (int i = 0; i < TB; ++i)
    if ((tid >> i) & 1) g_tid |= int64_t(1) << A.lpos[i + 1];

#pragma unroll
  for (int it = 0; it < (NR >> 1); ++it) {
    int64_t g_it = 0;
#pragma unroll
    for (int i = 0; i < R - 1; ++i)
      if ((it >> i) & 1) g_it |= int64_t(1) << A.lpos[TB + 1 + i];
    const int64_t gi = base | g_tid | g_it;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
    if (gi < total) {
      a = reinterpret_cast<const float4 *>(A.psi)[gi >> 1];
      b = reinterpret_cast<const float4 *>(A.lam)[gi >> 1];
    }
    const int l = swz(2 * (it * T + tid));
    *reinterpret_cast<float4 *>(&tp[l]) = a;
    *reinterpret_cast<float4 *>(&tl[l]) = b;
  }
  __syncthreads();

  float2 vp[NR], vl[NR];
  const int32_t *w = A.words;
  const float2 *cs = reinterpret_cast<const float2 *>(A.cs);
  for (int s = 0; s < A.nsub; ++s) {
    const int nops = w[0];
    int gl = 0;
    int64_t gbase = base;
#pragma unroll
    for (int i = 0; i < TB; ++i) {
      if ((tid >> i) & 1) {
        gl |= 1 << w[kHdrG + i];
        gbase |= int64_t(1) << w[kHdrGG + i];
      }
    }
    const int64_t sample = MULTI ? (gbase >> n) : block_sample;
    const int64_t csample = sample < A.B ? sample : 0;
    const int lb = swz(gl);
    int sw[R];
#pragma unroll
    for (int r = 0; r < R; ++r) sw[r] = w[kHdrSwz + r];
    auto addr = [&](int j) {
      int a = lb;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if ((j >> r) & 1) a ^= sw[r];
      return a;
    };
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      vp[j] = tp[addr(j)];
      vl[j] = tl[addr(j)];
    }

    const int32_t *op = w + kHdr;
    for (int o = 0; o < nops; ++o, op += 4) {
      const int code = op[0];
      const int a = op[1], b = op[2];
      const int kind = code & 0xff;
      if (kind == QS_OP_CNOT) {
        cnot_dispatch<R>(vp, b, a);
        cnot_dispatch<R>(vl, b, a);
        continue;
      }
      if (kind == QS_OP_CNOT_UNIFORM) {
        if ((gbase >> b) & 1) {
          swap_dispatch<R>(vp, a);
          swap_dispatch<R>(vl, a);
        }
        continue;
      }
      const int64_t idx = op[3] + ((code & QS_OP_PER_SAMPLE) ? csample : 0);
      const float2 c_s = cs[idx];
      float g;
      if (kind == QS_OP_RX) {
        g = im_dispatch<QS_RX, R>(vl, vp, a);
        rot_dispatch<QS_RX, R>(vp, a, c_s.x, -c_s.y);
        rot_dispatch<QS_RX, R>(vl, a, c_s.x, -c_s.y);
      } else if (kind == QS_OP_RY) {
        g = im_dispatch<QS_RY, R>(vl, vp, a);
        rot_dispatch<QS_RY, R>(vp, a, c_s.x, -c_s.y);
        rot_dispatch<QS_RY, R>(vl, a, c_s.x, -c_s.y);
      } else if (kind == QS_OP_RZ) {
        g = im_dispatch<QS_RZ, R>(vl, vp, a);
        rot_dispatch<QS_RZ, R>(vp, a, c_s.x, -c_s.y);
        rot_dispatch<QS_RZ, R>(vl, a, c_s.x, -c_s.y);
      } else {  // QS_OP_RZ_UNIFORM on global position a (constant per thread)
        const bool one = (gbase >> a) & 1;
        float acc = 0.f;
#pragma unroll
        for (int j = 0; j < NR; ++j) acc += vl[j].x * vp[j].y - vl[j].y * vp[j].x;
        g = one ? -acc : acc;
        const float c = c_s.x;
        const float sn = one ? c_s.y : -c_s.y;  // inverse phase
#pragma unroll
        for (int j = 0; j < NR; ++j) {
          float2 x = vp[j];
          vp[j] = make_float2(c * x.x + sn * x.y, c * x.y - sn * x.x);
          x = vl[j];
          vl[j] = make_float2(c * x.x + sn * x.y, c * x.y - sn * x.x);
        }
      }
      for (int off = 1; off < (1 << wlog); off <<= 1) g += __shfl_xor(g, off, 64);
      if (leader) A.gpart[(int64_t)b * A.NS + u] = g;
    }

#pragma unroll
    for (int j = 0; j < NR; ++j) {
      tp[addr(j)] = vp[j];
      tl[addr(j)] = vl[j];
    }
    __syncthreads();
    w += kHdr + 4 * nops;
  }

#pragma unroll
  for (int it = 0; it < (NR >> 1); ++it) {
    int64_t g_it = 0;
#pragma unroll
    for (int i = 0; i < R - 1; ++i)
      if ((it >> i) & 1) g_it |= int64_t(1) << A.lpos[TB + 1 + i];
    const int64_t gi = base | g_tid | g_it;
    const int l = swz(2 * (it * T + tid));
    if (gi < total) {
      reinterpret_cast<float4 *>(A.psi)[gi >> 1] = *reinterpret_cast<const float4 *>(&tp[l]);
      reinterpret_cast<float4 *>(A.lam)[gi >> 1] = *reinterpret_cast<const float4 *>(&tl[l]);
    }
  }
}

// out[row, b] = 1/2 sum of the row's slots of sample b; tot[row] = sum over b.
__global__ void __launch_bounds__(256)
k_adjoint_finish(float *out, float *tot, const float *gpart, int64_t B, int64_t NS, int64_t sps) {
  __shared__ double red[4];
  const int64_t row = blockIdx.x;
  const float *g = gpart + row * NS;
  double all = 0.0;
  for (int64_t b = threadIdx.x; b < B; b += 256) {
    double acc = 0.0;
    for (int64_t k = 0; k < sps; ++k) acc += g[b * sps + k];
    out[row * B + b] = (float)(0.5 * acc);
    all += 0.5 * acc;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) all += __shfl_xor(all, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = all;
  __syncthreads();
  if (threadIdx.x == 0) tot[row] = (float)(red[0] + red[1] + red[2] + red[3]);
}

}  // namespace qs

using namespace qs;

static void adj_geometry(int64_t B, int n, int R, int TB, int64_t *ntiles, int64_t *NS, int64_t *sps) {
  const int K = R + TB;
  const bool multi = n < K;
  const int seglog = multi ? (n - R) : TB;
  const int wlog = seglog < 6 ? seglog : 6;
  *ntiles = multi ? (B + (int64_t(1) << (K - n)) - 1) >> (K - n) : (B << (n - K));
  *NS = *ntiles * (int64_t(1) << (TB - wlog));
  *sps = multi ? (int64_t(1) << (seglog - wlog)) : ((int64_t(1) << (n - K)) << (TB - wlog));
}

template <int R, int TB>
static int launch_adj(const AdjArgs &A, bool multi, int64_t ntiles, hipStream_t st) {
  if (multi)
    hipLaunchKernelGGL((k_adjoint_pass<R, TB, true>), dim3((unsigned)ntiles), dim3(1 << TB), 0, st, A);
  else
    hipLaunchKernelGGL((k_adjoint_pass<R, TB, false>), dim3((unsigned)ntiles), dim3(1 << TB), 0, st, A);
  return launch_status();
}

extern "C" {

int64_t qs_adjoint_work_bytes(int64_t B, int n, int R, int TB, int rows) {
  if (B < 1 || n < 1 || n > QS_MAX_QUBITS || rows < 0 || !qs_fused_supported(R, TB) || R > n)
    return -1;
  int64_t ntiles, NS, sps;
  adj_geometry(B, n, R, TB, &ntiles, &NS, &sps);
  return (int64_t)(rows > 0 ? rows : 1) * NS * (int64_t)sizeof(float);
}

int qs_adjoint_pass(float *psi, float *lam, int64_t B, int n, int R, int TB, const int32_t *lpos,
                    const int32_t *words, int nsub, const float *cs, float *gpart, int rows,
                    void *stream) {
  if (B < 1 || n < 1 || n > QS_MAX_QUBITS || !psi || !lam || !lpos || !words || nsub < 1 ||
      !gpart || rows < 0 || !qs_fused_supported(R, TB) || R > n)
    return QS_EINVAL;
  const int K = R + TB;
  AdjArgs A;
  A.psi = psi;
  A.lam = lam;
  A.cs = cs;
  A.gpart = gpart;
  A.words = words;
  A.B = B;
  A.n = n;
  A.nsub = nsub;
  int64_t ntiles, NS, sps;
  adj_geometry(B, n, R, TB, &ntiles, &NS, &sps);
  A.NS = NS;
  for (int j = 0; j < QS_MAX_LOCAL_BITS; ++j) A.lpos[j] = 0;
  for (int j = 0; j < QS_MAX_QUBITS; ++j) A.npos[j] = 0;
  const bool multi = n < K;
  if (multi) {
    for (int j = 0; j < K; ++j) A.lpos[j] = j;
    A.nn = 0;
  } else {
    int64_t seen = 0;
    for (int j = 0; j < K; ++j) {
      if (lpos[j] < 0 || lpos[j] >= n || ((seen >> lpos[j]) & 1)) return QS_EINVAL;
      if (j > 0 && lpos[j] <= lpos[j - 1]) return QS_EINVAL;
      seen |= int64_t(1) << lpos[j];
      A.lpos[j] = lpos[j];
    }
    if (A.lpos[0] != 0) return QS_EINVAL;
    int c = 0;
    for (int p = 0; p < n; ++p)
      if (!((seen >> p) & 1)) A.npos[c++] = p;
    A.nn = c;
  }
  if (ntiles > 0x7fffffff) return QS_EUNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  if (TB == 8) {
    switch (R) {
      case 1: return launch_adj<1, 8>(A, multi, ntiles, st);
      case 2: return launch_adj<2, 8>(A, multi, ntiles, st);
      case 3: return launch_adj<3, 8>(A, multi, ntiles, st);
      case 4: return launch_adj<4, 8>(A, multi, ntiles, st);
      case 5: return launch_adj<5, 8>(A, multi, ntiles, st);
    }
  }
  if (TB == 7 && R == 5) return launch_adj<5, 7>(A, multi, ntiles, st);
  if (TB == 6 && R == 5) return launch_adj<5, 6>(A, multi, ntiles, st);
  return QS_EUNSUPPORTED;
}

int qs_adjoint_finish(float *out, float *tot, const float *gpart, int rows, int64_t B, int n,
                      int R, int TB, void *stream) {
  if (B < 1 || n < 1 || n > QS_MAX_QUBITS || rows < 0 || !qs_fused_supported(R, TB) || R > n)
    return QS_EINVAL;
  if (rows == 0) return QS_OK;
  if (!out || !tot || !gpart) return QS_EINVAL;
  int64_t ntiles, NS, sps;
  adj_geometry(B, n, R, TB, &ntiles, &NS, &sps);
  hipLaunchKernelGGL(k_adjoint_finish, dim3((unsigned)rows), dim3(256), 0, (hipStream_t)stream, out,
                     tot, gpart, B, NS, sps);
  return launch_status();
}

}  // extern "C"
