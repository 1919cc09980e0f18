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
#include <string.h>

#include "qs_pass_common.h"

namespace qs {

struct AdjArgs {
  float *psi;
  float *lam;
  const int32_t *cs;    // (cos, sin) float pairs, read as int32 bits
  const int32_t *utab;  // fused per-bit unitaries (un-applied as U^dagger), int32 bits
  const int32_t *qtab;  // gradient generators of their constituents, int32 bits
  float *gpart;
  const int32_t *words;
  const float *seed_w;  // non-null: lam = 2 psi sum_q w[b, q] [bit n-1-q = 0], (B, n) row-major
  int64_t B;
  int64_t NS;  // reduction slots per gradient row
  int n;
  int nsub;
  int nn;
  int32_t lpos[QS_MAX_LOCAL_BITS];
  int32_t npos[QS_MAX_QUBITS];
};

template <int R, int TB, bool MULTI, class AP>
__device__ __forceinline__ void adjoint_tile(AP A, const int64_t t, float2 *tp, float2 *tl) {
  constexpr int T = 1 << TB;
  constexpr int K = R + TB;
  constexpr int NR = 1 << R;

  const int tid = threadIdx.x;
  const int n = A->n;
  const int64_t total = A->B << n;
  int64_t base;
  if (MULTI) {
    base = t << K;
  } else {
    const int64_t tlo = t & ((int64_t(1) << A->nn) - 1);
    base = (t >> A->nn) << n;
    for (int i = 0; i < A->nn; ++i)
      if ((tlo >> i) & 1) base |= int64_t(1) << A->npos[i];
  }
  const int64_t block_sample = base >> n;
  const int seglog = MULTI ? (n - R) : TB;
  const int wlog = seglog < 6 ? seglog : 6;
  const int64_t u = t * (T >> wlog) + (tid >> wlog);  // reduction slot of this thread's segment
  const bool leader = (tid & ((1 << wlog) - 1)) == 0;

  int64_t g_tid = 0;
#pragma unroll
  for (int i = 0; i < TB; ++i)
    if ((tid >> i) & 1) g_tid |= int64_t(1) << A->lpos[i + 1];

#pragma unroll
  for (int it = 0; it < (NR >> 1); ++it) {
    int64_t g_it = 0;
#pragma unroll
    for (int i = 0; i < R - 1; ++i)
      if ((it >> i) & 1) g_it |= int64_t(1) << A->lpos[TB + 1 + i];
    const int64_t gi = base | g_tid | g_it;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
    if (gi < total) {
      a = load_g4(A->psi, gi >> 1);
      if (A->seed_w) {
        // co-state seed of MeasureProbability for the pair (gi, gi + 1)
        const kptr<float> wb = to_const(A->seed_w) + (gi >> n) * n;
        float w0 = 0.f, w1 = 0.f;
        for (int pos = 1; pos < n; ++pos)
          if (!((gi >> pos) & 1)) w0 += wb[n - 1 - pos];
        w1 = w0;
        w0 += wb[n - 1];  // position 0 is 0 for gi (even) and 1 for gi + 1
        b = make_float4(2.f * w0 * a.x, 2.f * w0 * a.y, 2.f * w1 * a.z, 2.f * w1 * a.w);
      } else {
        b = load_g4(A->lam, gi >> 1);
      }
    }
    const int sl = swz(2 * (it * T + tid));
    if (sl & 1) {
      a = make_float4(a.z, a.w, a.x, a.y);
      b = make_float4(b.z, b.w, b.x, b.y);
    }
    *reinterpret_cast<float4 *>(&tp[sl & ~1]) = a;
    *reinterpret_cast<float4 *>(&tl[sl & ~1]) = b;
  }
  __syncthreads();

  float2 vp[NR], vl[NR];
  kptr<int32_t> w = to_const(A->words);
  for (int s = 0; s < A->nsub; ++s) {
    const int nseg = w[0], nphase = w[6];
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
    const int64_t csample = sample < A->B ? sample : 0;
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

    // A segment's sums go to its reduction slot u.  Whole-wave segments sum
    // four rows at a time with DPP and lane swaps (wave_sum4; lanes 0..3 store
    // them); narrower segments (several samples per wave) reduce by shuffles.
    const bool wave_slots = wlog == 6;
    const gptr<float> gp = to_global(A->gpart);
    auto put1 = [&](float g, int row) {
      for (int off = 1; off < (1 << wlog); off <<= 1) g += __shfl_xor(g, off, 64);
      if (leader && row >= 0) gp[(int64_t)row * A->NS + u] = g;
    };
    auto store4 = [&](float g, const int (&rw)[4]) {  // g: a wave_sum4 result
      const int ln = tid & 63;
      if (ln < 4) {
        const int v = wave_sum4_index(ln);
        const int row = v == 0 ? rw[0] : v == 1 ? rw[1] : v == 2 ? rw[2] : rw[3];
        if (row >= 0) gp[(int64_t)row * A->NS + u] = g;
      }
    };
    const kptr<int32_t> ph = w + kHdr;
    kptr<int32_t> op = ph + 4 * nphase;
    if (nphase) {
      // every phase entry commutes with the whole sub-pass: its gradient is read
      // at the sub-pass output, where Im<lam|Z_p|psi> = (+/-) sum Im(conj(lam) psi)
      float S = 0.f;
#pragma unroll
      for (int j = 0; j < NR; ++j) S += vl[j].x * vp[j].y - vl[j].y * vp[j].x;
      if (wave_slots) {
        for (int k0 = 0; k0 < nphase; k0 += 4) {
          float tv[4];
          int rw[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const bool ok = k0 + c < nphase;
            const int k = ok ? k0 + c : k0;
            tv[c] = ok ? (((gbase >> ph[4 * k]) & 1) ? -S : S) : 0.f;
            rw[c] = ok ? ph[4 * k + 3] : -1;
          }
          store4(wave_sum4(tv[0], tv[1], tv[2], tv[3]), rw);
        }
      } else {
        for (int k = 0; k < nphase; ++k) put1(((gbase >> ph[4 * k]) & 1) ? -S : S, ph[4 * k + 3]);
      }
      const float2 z = phase_product(ph, nphase, to_const(A->cs), csample, gbase, true);
#pragma unroll
      for (int j = 0; j < NR; ++j) {
        vp[j] = cmul(z, vp[j]);
        vl[j] = cmul(z, vl[j]);
      }
    }
    for (int sg = 0; sg < nseg; ++sg) {
      const int head = op[0];
      const int umask = head & 0xff, ncnot = (head >> 8) & 0xff;
      const int psmask = (head >> 16) & 0xff, probemask = (head >> 24) & 0x7f;
      ++op;
      for (int c = 0; c < ncnot; ++c) {
        run_cnot<R>(vp, op[c], gbase);
        run_cnot<R>(vl, op[c], gbase);
      }
      op += ncnot;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if ((umask >> r) & 1) {
          const int nconst = op[1];
          const bool ps = (psmask >> r) & 1;
          const bool probe = (probemask >> r) & 1;
          const int64_t var = ps ? csample : 0;
          const int64_t span = ps ? A->B : 1;
          const kptr<int32_t> qt = to_const(A->qtab);
          const M8 uu = load_m8(to_const(A->utab) + 8 * (op[0] + var));  // a probe's entry is the identity
          const H4 m = herm_density_at<R>(vp, vl, r);  // at the gate's output
          // gradient rows k0, k0 + 1, ... (four at a time): traces and the wave sum
          auto sum4 = [&](int k0, int (&rw)[4]) {
            float tv[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              const bool ok = k0 + c < nconst;
              const int k = ok ? k0 + c : k0;
              const float tr = herm_trace(qt + 8 * (op[2] + k * span + var), m);
              tv[c] = ok ? tr : 0.f;
              rw[c] = ok ? op[3 + k] : -1;
            }
            return wave_sum4(tv[0], tv[1], tv[2], tv[3]);
          };
          if (wave_slots && !probe && !(MULTI && ps)) {
            // the first four rows' traces and wave sum sit in one block with the
            // U^dagger matrix-core chains, so the scheduler can interleave them
            int rw[4];
            const float g = sum4(0, rw);
            const MU ud = mfma_coeffs<true>(uu);
            apply_mfma_at<R>(vp, r, ud);
            apply_mfma_at<R>(vl, r, ud);
            store4(g, rw);
            for (int k0 = 4; k0 < nconst; k0 += 4) {
              const float g2 = sum4(k0, rw);
              store4(g2, rw);
            }
          } else {
            if (wave_slots) {
              for (int k0 = 0; k0 < nconst; k0 += 4) {
                int rw[4];
                const float g = sum4(k0, rw);
                store4(g, rw);
              }
            } else {
              for (int k = 0; k < nconst; ++k) put1(herm_trace(qt + 8 * (op[2] + k * span + var), m), op[3 + k]);
            }
            if (!probe) {
              if (MULTI && ps) {  // may differ inside a 4-lane matrix block
                apply_valu_at<R>(vp, r, dagger(uu));
                apply_valu_at<R>(vl, r, dagger(uu));
              } else {
                const MU ud = mfma_coeffs<true>(uu);
                apply_mfma_at<R>(vp, r, ud);
                apply_mfma_at<R>(vl, r, ud);
              }
            }
          }
          op += 3 + nconst;
        }
      }
    }

#pragma unroll
    for (int j = 0; j < NR; ++j) {
      tp[addr(j)] = vp[j];
      tl[addr(j)] = vl[j];
    }
    __syncthreads();
    w = op;
  }

#pragma unroll
  for (int it = 0; it < (NR >> 1); ++it) {
    int64_t g_it = 0;
#pragma unroll
    for (int i = 0; i < R - 1; ++i)
      if ((it >> i) & 1) g_it |= int64_t(1) << A->lpos[TB + 1 + i];
    const int64_t gi = base | g_tid | g_it;
    const int sl = swz(2 * (it * T + tid));
    float4 a = *reinterpret_cast<const float4 *>(&tp[sl & ~1]);
    float4 b = *reinterpret_cast<const float4 *>(&tl[sl & ~1]);
    if (sl & 1) {
      a = make_float4(a.z, a.w, a.x, a.y);
      b = make_float4(b.z, b.w, b.x, b.y);
    }
    if (gi < total) {
      store_g4(A->psi, gi >> 1, a);
      store_g4(A->lam, gi >> 1, b);
    }
  }
}

template <int R, int TB, bool MULTI>
__global__ void __launch_bounds__(1 << TB) k_adjoint_pass(AdjArgs A) {
  __shared__ __align__(16) float2 tp[1 << (R + TB)];
  __shared__ __align__(16) float2 tl[1 << (R + TB)];
  adjoint_tile<R, TB, MULTI>(&A, blockIdx.x, tp, tl);
}

template <int R, int TB>
__global__ void __launch_bounds__(1 << TB)
k_adjoint_items(const AdjArgs *descs, const int32_t *items, int nitems) {
  __shared__ __align__(16) float2 tp[1 << (R + TB)];
  __shared__ __align__(16) float2 tl[1 << (R + TB)];
  const kptr<AdjArgs> D = to_const(descs);
  const kptr<int32_t> I = to_const(items);
  for (int i = blockIdx.x; i < nitems; i += gridDim.x) {
    const kptr<AdjArgs> A = D + I[2 * i];
    const int64_t t = I[2 * i + 1];
    if (A->n < R + TB)
      adjoint_tile<R, TB, true>(A, t, tp, tl);
    else
      adjoint_tile<R, TB, false>(A, t, tp, tl);
    __syncthreads();
  }
}

// out[row, b] = 1/2 sum of the row's slots of sample b; tot[row] = sum over b.
struct AdjFinishArgs {
  float *out;
  float *tot;
  const float *gpart;
  int64_t B;
  int64_t NS;
  int64_t sps;
  int rows;
};

// The slots of sample b are contiguous (sps of them, a power of two): each wave reads 64 consecutive
// slots per load and sums them in segments of sps lanes with shuffles (in double), so the loads are
// coalesced; with more than 64 slots per sample a wave takes one sample at a time.  The loop trip
// counts are uniform per wave, so every lane takes part in every shuffle.
__device__ __forceinline__ void adjoint_finish_row(const AdjFinishArgs &F, int64_t row, double *red) {
  const float *g = F.gpart + row * F.NS;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double all = 0.0;
  if (F.sps <= 64) {
    const int64_t total = F.B * F.sps;
    const int seg = (int)F.sps;
    // eight chunks (wave, wave + 4, ... of 64 slots) per step, their loads in flight together
    // before their shuffle sums
    for (int64_t c0 = (int64_t)wave * 64; c0 < total; c0 += 2048) {
      double acc[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int64_t i = c0 + u * 256 + lane;
        acc[u] = i < total ? (double)g[i] : 0.0;
      }
      for (int off = 1; off < seg; off <<= 1) {
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] += __shfl_xor(acc[u], off, 64);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int64_t i = c0 + u * 256 + lane;
        if ((lane & (seg - 1)) == 0 && i < total) {
          F.out[row * F.B + i / seg] = (float)(0.5 * acc[u]);
          all += 0.5 * acc[u];
        }
      }
    }
  } else {
    for (int64_t b = wave; b < F.B; b += 4) {
      const float *gb = g + b * F.sps;
      double acc = 0.0;
      for (int64_t k = lane; k < F.sps; k += 64) acc += (double)gb[k];
      for (int off = 1; off < 64; off <<= 1) acc += __shfl_xor(acc, off, 64);
      if (lane == 0) {
        F.out[row * F.B + b] = (float)(0.5 * acc);
        all += 0.5 * acc;
      }
    }
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) all += __shfl_xor(all, off, 64);
  __syncthreads();
  if (lane == 0) red[wave] = all;
  __syncthreads();
  if (threadIdx.x == 0) F.tot[row] = (float)(red[0] + red[1] + red[2] + red[3]);
}

__global__ void __launch_bounds__(256) k_adjoint_finish(AdjFinishArgs F) {
  __shared__ double red[4];
  adjoint_finish_row(F, blockIdx.x, red);
}

// One wave per gradient row (the rows of a batch of circuits are many and
// short: a circuit of batch 1 has at most a few hundred slots per row).  The
// same sums as adjoint_finish_row, in double, with no block barrier between
// rows, so every wave of the launch keeps its own rows' loads in flight.
__device__ __forceinline__ void adjoint_finish_row_wave(kptr<AdjFinishArgs> F, int64_t row, int lane) {
  const gptr<const float> g = to_global(F->gpart) + row * F->NS;
  const gptr<float> out = to_global(F->out) + row * F->B;
  const int64_t B = F->B, sps = F->sps;
  double all = 0.0;
  if (sps <= 64) {
    const int64_t total = B * sps;
    const int seg = (int)sps;
    for (int64_t c0 = 0; c0 < total; c0 += 512) {
      double acc[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int64_t i = c0 + u * 64 + lane;
        acc[u] = i < total ? (double)g[i] : 0.0;
      }
      for (int off = 1; off < seg; off <<= 1) {
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] += __shfl_xor(acc[u], off, 64);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int64_t i = c0 + u * 64 + lane;
        if ((lane & (seg - 1)) == 0 && i < total) {
          out[i / seg] = (float)(0.5 * acc[u]);
          all += 0.5 * acc[u];
        }
      }
    }
  } else {
    for (int64_t b = 0; b < B; ++b) {
      const gptr<const float> gb = g + b * sps;
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      for (int64_t k = lane; k < sps; k += 256) {
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] += k + 64 * u < sps ? (double)gb[k + 64 * u] : 0.0;
      }
      double a = (acc[0] + acc[1]) + (acc[2] + acc[3]);
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) a += __shfl_xor(a, off, 64);
      if (lane == 0) {
        out[b] = (float)(0.5 * a);
        all += 0.5 * a;
      }
    }
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) all += __shfl_xor(all, off, 64);
  if (lane == 0) to_global(F->tot)[row] = (float)all;
}

// Block (circuit, row block): wave w of row block y takes rows 4 y + w, then
// every 4 gridDim.y rows after it.
__global__ void __launch_bounds__(256) k_adjoint_finish_items(const AdjFinishArgs *descs) {
  const kptr<AdjFinishArgs> F = to_const(descs) + blockIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  for (int64_t row = 4 * (int64_t)blockIdx.y + wave; row < F->rows; row += 4 * (int64_t)gridDim.y)
    adjoint_finish_row_wave(F, row, lane);
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
  if (B < 1 || n < 1 || n > QS_MAX_QUBITS || rows < 0 || !qs_layout_supported(R, TB) || R > n)
    return -1;
  int64_t ntiles, NS, sps;
  adj_geometry(B, n, R, TB, &ntiles, &NS, &sps);
  return (int64_t)(rows > 0 ? rows : 1) * NS * (int64_t)sizeof(float);
}

}  // extern "C"

static int fill_adj_args(AdjArgs &A, int64_t *ntiles, float *psi, float *lam, int64_t B, int n, int R,
                         int TB, const int32_t *lpos, const int32_t *words, int nsub, const float *cs,
                         const float *utab, const float *qtab, float *gpart, int rows,
                         const float *seed_w) {
  if (B < 1 || n < 1 || n > QS_MAX_QUBITS || !psi || !lam || !lpos || !words || nsub < 1 ||
      !gpart || rows < 0 || !qs_fused_supported(R, TB) || R > n)
    return QS_EINVAL;
  const int K = R + TB;
  A.psi = psi;
  A.lam = lam;
  A.cs = reinterpret_cast<const int32_t *>(cs);
  A.utab = reinterpret_cast<const int32_t *>(utab);
  A.qtab = reinterpret_cast<const int32_t *>(qtab);
  A.gpart = gpart;
  A.words = words;
  A.seed_w = seed_w;
  A.B = B;
  A.n = n;
  A.nsub = nsub;
  int64_t NS, sps;
  adj_geometry(B, n, R, TB, ntiles, &NS, &sps);
  A.NS = NS;
  for (int j = 0; j < QS_MAX_LOCAL_BITS; ++j) A.lpos[j] = 0;
  for (int j = 0; j < QS_MAX_QUBITS; ++j) A.npos[j] = 0;
  if (n < K) {
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
  return QS_OK;
}

extern "C" {

int qs_adjoint_pass(float *psi, float *lam, int64_t B, int n, int R, int TB, const int32_t *lpos,
                    const int32_t *words, int nsub, const float *cs, const float *utab,
                    const float *qtab, float *gpart, int rows, const float *seed_w, void *stream) {
  AdjArgs A;
  int64_t ntiles;
  int st0 = fill_adj_args(A, &ntiles, psi, lam, B, n, R, TB, lpos, words, nsub, cs, utab, qtab, gpart,
                          rows, seed_w);
  if (st0) return st0;
  if (ntiles > 0x7fffffff) return QS_EUNSUPPORTED;
  const bool multi = n < R + TB;
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
  if (TB == 7 && R == 4) return launch_adj<4, 7>(A, multi, ntiles, st);
  if (TB == 6 && R == 5) return launch_adj<5, 6>(A, multi, ntiles, st);
  return QS_EUNSUPPORTED;
}

int64_t qs_adjoint_desc_bytes(void) { return (int64_t)sizeof(AdjArgs); }

int qs_adjoint_desc(void *desc, int64_t *ntiles, float *psi, float *lam, int64_t B, int n, int R,
                    int TB, const int32_t *lpos, const int32_t *words, int nsub, const float *cs,
                    const float *utab, const float *qtab, float *gpart, int rows, const float *seed_w) {
  if (!desc || !ntiles) return QS_EINVAL;
  AdjArgs A;
  int st0 = fill_adj_args(A, ntiles, psi, lam, B, n, R, TB, lpos, words, nsub, cs, utab, qtab, gpart,
                          rows, seed_w);
  if (st0) return st0;
  memcpy(desc, &A, sizeof(AdjArgs));
  return QS_OK;
}

int qs_items_adjoint(const void *descs, const int32_t *items, int nitems, int R, int TB, int grid,
                     void *stream) {
  if (!descs || !items || nitems < 0 || grid < 1) return QS_EINVAL;
  if (nitems == 0) return QS_OK;
  hipStream_t st = (hipStream_t)stream;
  const AdjArgs *d = (const AdjArgs *)descs;
  const unsigned g = (unsigned)(grid < nitems ? grid : nitems);
  if (R == 4 && TB == 7) {
    hipLaunchKernelGGL((k_adjoint_items<4, 7>), dim3(g), dim3(128), 0, st, d, items, nitems);
    return launch_status();
  }
  if (R == 4 && TB == 8) {
    hipLaunchKernelGGL((k_adjoint_items<4, 8>), dim3(g), dim3(256), 0, st, d, items, nitems);
    return launch_status();
  }
  if (R == 5 && TB == 6) {
    hipLaunchKernelGGL((k_adjoint_items<5, 6>), dim3(g), dim3(64), 0, st, d, items, nitems);
    return launch_status();
  }
  if (R == 5 && TB == 8) {
    hipLaunchKernelGGL((k_adjoint_items<5, 8>), dim3(g), dim3(256), 0, st, d, items, nitems);
    return launch_status();
  }
  return QS_EUNSUPPORTED;
}

static int fill_adj_finish(AdjFinishArgs &F, float *out, float *tot, const float *gpart, int rows,
                           int64_t B, int n, int R, int TB) {
  if (B < 1 || n < 1 || n > QS_MAX_QUBITS || rows < 0 || !qs_layout_supported(R, TB) || R > n)
    return QS_EINVAL;
  if (rows > 0 && (!out || !tot || !gpart)) return QS_EINVAL;
  int64_t ntiles;
  F.out = out;
  F.tot = tot;
  F.gpart = gpart;
  F.B = B;
  F.rows = rows;
  adj_geometry(B, n, R, TB, &ntiles, &F.NS, &F.sps);
  return QS_OK;
}

int qs_adjoint_finish(float *out, float *tot, const float *gpart, int rows, int64_t B, int n,
                      int R, int TB, void *stream) {
  AdjFinishArgs F;
  int st0 = fill_adj_finish(F, out, tot, gpart, rows, B, n, R, TB);
  if (st0) return st0;
  if (rows == 0) return QS_OK;
  hipLaunchKernelGGL(k_adjoint_finish, dim3((unsigned)rows), dim3(256), 0, (hipStream_t)stream, F);
  return launch_status();
}

int64_t qs_adjoint_finish_desc_bytes(void) { return (int64_t)sizeof(AdjFinishArgs); }

int qs_adjoint_finish_desc(void *desc, float *out, float *tot, const float *gpart, int rows,
                           int64_t B, int n, int R, int TB) {
  if (!desc) return QS_EINVAL;
  AdjFinishArgs F;
  int st0 = fill_adj_finish(F, out, tot, gpart, rows, B, n, R, TB);
  if (st0) return st0;
  memcpy(desc, &F, sizeof(AdjFinishArgs));
  return QS_OK;
}

int qs_items_adjoint_finish(const void *descs, int ndesc, int row_blocks, void *stream) {
  if (!descs || ndesc < 0 || row_blocks < 1) return QS_EINVAL;
  if (ndesc == 0) return QS_OK;
  hipLaunchKernelGGL(k_adjoint_finish_items, dim3((unsigned)ndesc, (unsigned)row_blocks), dim3(256),
                     0, (hipStream_t)stream, (const AdjFinishArgs *)descs);
  return launch_status();
}

}  // extern "C"
