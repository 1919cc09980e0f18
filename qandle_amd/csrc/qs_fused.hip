// Fused pass kernel for gfx950: many gates per HBM round trip.
//
// One workgroup owns one tile: the 2^K amplitudes whose non-local index bits
// are fixed (K = R + TB local bits; the low local bits are the lowest global
// positions so loads/stores are contiguous 16-byte vectors).  The tile is
// staged in LDS.  A pass is a list of sub-passes; in each sub-pass every
// thread holds the 2^R amplitudes that differ in the sub-pass's R "register"
// bits and applies that sub-pass's gates entirely in VGPRs (register indices
// are compile-time, the dispatch over bit positions is a uniform branch).
// Gates whose second bit is not a register bit are still possible when that
// bit is only a CNOT control or an RZ: the bit's value is then constant per
// thread.  The first pass of a circuit generates the AngleEmbedding product
// state in registers instead of loading; the last pass can fold
// MeasureProbability into a deterministic reduction.
//
// Replaces the executor loop of UnsplittedCircuit.forward
// (src/qandle/qcircuit.py:163-174) and the dense per-gate `state @ M` it runs
// (src/qandle/operators.py:289-298, 557-558), the embedding's matrix chain
// (src/qandle/embeddings.py:148-164) and MeasureProbability's masked reduce
// (src/qandle/measurements.py:120-123).
#include <string.h>

#include "qs_pass_common.h"

namespace qs {

struct PassArgs {
  float *state;
  const float *src;
  const int32_t *cs;    // (cos, sin) float pairs, read as int32 bits
  const int32_t *utab;  // fused per-bit unitaries, 8 floats each, read as int32 bits
  const float *efac;
  float *mpart;
  const int32_t *words;
  int64_t B;
  int n;
  int nsub;
  int store;
  int multi;  // tile spans 2^(K-n) whole samples (n < K)
  int nn;     // number of non-local qubit positions (n > K)
  int32_t lpos[QS_MAX_LOCAL_BITS];
  int32_t npos[QS_MAX_QUBITS];
};

// Global index of tile t's first amplitude (its non-local bits).
template <class AP>
__device__ __forceinline__ int64_t tile_base(AP A, int64_t t, int K, bool multi) {
  if (multi) return t << K;
  const int64_t tl = t & ((int64_t(1) << A->nn) - 1);
  int64_t base = (t >> A->nn) << A->n;
  for (int i = 0; i < A->nn; ++i)
    if ((tl >> i) & 1) base |= int64_t(1) << A->npos[i];
  return base;
}

// Offset of this thread's first load/store slot: thread bit i is local bit
// i + 1 (local bit 0 is the amplitude pair inside one 16-byte vector).
template <int TB, class AP>
__device__ __forceinline__ int64_t thread_offset(AP A, int tid) {
  int64_t g = 0;
#pragma unroll
  for (int i = 0; i < TB; ++i)
    if ((tid >> i) & 1) g |= int64_t(1) << A->lpos[i + 1];
  return g;
}

// Issues the HBM loads of one tile into registers (slot it of this thread =
// amplitude pair it * T + tid of the tile, in local-index order).  The loads
// stay in flight until the registers are staged into LDS.
template <int R, int TB, class AP>
__device__ __forceinline__ void fetch_tile(AP A, int64_t base, int64_t g_tid,
                                           float4 (&pre)[1 << (R - 1)]) {
  const int64_t total = A->B << A->n;
#pragma unroll
  for (int it = 0; it < (1 << (R - 1)); ++it) {
    int64_t g_it = 0;
#pragma unroll
    for (int i = 0; i < R - 1; ++i)
      if ((it >> i) & 1) g_it |= int64_t(1) << A->lpos[TB + 1 + i];
    const int64_t gi = base | g_tid | g_it;
    pre[it] = gi < total ? load_g4(A->src, gi >> 1) : make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// Writes fetched registers to their swizzled LDS slots (an amplitude pair
// shares one 16-byte slot, swapped when bit 0 of its swizzled index is set).
template <int R, int TB>
__device__ __forceinline__ void stage_tile(float2 *tile, const float4 (&pre)[1 << (R - 1)], int tid) {
#pragma unroll
  for (int it = 0; it < (1 << (R - 1)); ++it) {
    const int sl = swz(2 * (it * (1 << TB) + tid));
    float4 val = pre[it];
    if (sl & 1) val = make_float4(val.z, val.w, val.x, val.y);
    *reinterpret_cast<float4 *>(&tile[sl & ~1]) = val;
  }
}

// Sub-passes of one tile on its LDS copy (staged by the caller when the pass
// reads a state; generated in registers when it does not), then the store.
// `tile` (2^K float2, 16-byte aligned) and `red` ((T/64)*(K+1) floats) are
// the caller's LDS.
template <int R, int TB, bool MULTI, class AP>
__device__ __forceinline__ void pass_body(AP A, const int64_t t, const int64_t base,
                                          const int64_t g_tid, float2 *tile, float *red) {
  constexpr int T = 1 << TB;
  constexpr int K = R + TB;
  constexpr int NR = 1 << R;

  const int tid = threadIdx.x;
  const int n = A->n;
  const int64_t total = A->B << n;
  const int64_t block_sample = base >> n;

  float2 v[NR];
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
    const bool live = sample < A->B;
    const int lb = swz(gl);
    int sw[R];
#pragma unroll
    for (int r = 0; r < R; ++r) sw[r] = w[kHdrSwz + r];
    // LDS index of register j: swz is linear over XOR, so it is lb ^ (xor of the
    // swizzled strides of j's set bits)
    auto addr = [&](int j) {
      int a = lb;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if ((j >> r) & 1) a ^= sw[r];
      return a;
    };

    if (s == 0 && !A->src) {
      // product-state generator: (x)_pos f_pos(bit), per sample
      if (live) {
        const kptr<int32_t> E = to_const(reinterpret_cast<const int32_t *>(A->efac)) + sample * n * 4;
        int64_t rmask = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) rmask |= int64_t(1) << w[kHdrGP + r];
        float2 pre = make_float2(1.f, 0.f);
        for (int pos = 0; pos < n; ++pos)
          if (!((rmask >> pos) & 1)) pre = cmul(pre, load_k2(E, pos * 2 + ((gbase >> pos) & 1)));
        // doubling over the register bits; both loops have constant trip
        // counts so every v[] index is a compile-time constant
        v[0] = pre;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const float2 f0 = load_k2(E, w[kHdrGP + r] * 2), f1 = load_k2(E, w[kHdrGP + r] * 2 + 1);
#pragma unroll
          for (int j = 0; j < (1 << r); ++j) {
            v[j | (1 << r)] = cmul(v[j], f1);
            v[j] = cmul(v[j], f0);
          }
        }
      } else {
#pragma unroll
        for (int j = 0; j < NR; ++j) v[j] = make_float2(0.f, 0.f);
      }
    } else {
#pragma unroll
      for (int j = 0; j < NR; ++j) v[j] = tile[addr(j)];
    }

    // a padding sample past B (last tile of a multi-sample pass) holds zeros;
    // it reads sample 0's coefficients so every table access stays in bounds
    const int64_t csample = live ? sample : 0;
    const kptr<int32_t> ph = w + kHdr;
    kptr<int32_t> op = ph + 4 * nphase;
    for (int sg = 0; sg < nseg; ++sg) {
      const int head = op[0];
      const int umask = head & 0xff, ncnot = (head >> 8) & 0xff, psmask = (head >> 16) & 0xff;
      // every table read of the segment is issued before the first use
      M8 u[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int64_t ui = op[1 + r] + (((psmask >> r) & 1) ? csample : 0);
        u[r] = load_m8(to_const(A->utab) + 8 * ui);
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if ((umask >> r) & 1) {
          // per-sample unitaries of a multi-sample tile may differ inside a
          // 4-lane matrix block: those stay on the vector ALU
          if (MULTI && ((psmask >> r) & 1))
            apply_valu_at<R>(v, r, u[r]);
          else
            apply_mfma_at<R>(v, r, mfma_coeffs<false>(u[r]));
        }
      }
      op += 1 + R;
      for (int c = 0; c < ncnot; ++c) run_cnot<R>(v, op[c], gbase);
      op += ncnot;
    }
    if (nphase) {
      const float2 z = phase_product(ph, nphase, to_const(A->cs), csample, gbase, false);
#pragma unroll
      for (int j = 0; j < NR; ++j) v[j] = cmul(z, v[j]);
    }

    if (s == A->nsub - 1 && A->mpart) {
      // MeasureProbability partials: for every local bit, sum of |a|^2 over the
      // tile's amplitudes (of this sample) with that bit = 0; then the total.
      float p[NR];
      float tot = 0.f;
#pragma unroll
      for (int j = 0; j < NR; ++j) {
        p[j] = v[j].x * v[j].x + v[j].y * v[j].y;
        tot += p[j];
      }
      float c[K + 1];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float acc = 0.f;
#pragma unroll
        for (int j = 0; j < NR; ++j)
          if (!(j & (1 << r))) acc += p[j];
        c[r] = acc;
      }
#pragma unroll
      for (int i = 0; i < TB; ++i) c[R + i] = ((tid >> i) & 1) ? 0.f : tot;
      c[K] = tot;
      // reduce over the threads of one sample (a power-of-two tid segment)
      const int seglog = MULTI ? (n - R) : TB;
      const int wlog = seglog < 6 ? seglog : 6;
#pragma unroll
      for (int e = 0; e <= K; ++e) {
        float x = c[e];
        for (int off = 1; off < (1 << wlog); off <<= 1) x += __shfl_xor(x, off, 64);
        c[e] = x;
      }
      const int spt_log = MULTI ? (K - n) : 0;
      const gptr<float> mp = to_global(A->mpart) + (t << spt_log) * (K + 1);
      if (seglog <= 6) {
        if ((tid & ((1 << seglog) - 1)) == 0) {
          const gptr<float> o = mp + (tid >> seglog) * (K + 1);
#pragma unroll
          for (int r = 0; r < R; ++r) o[w[kHdrP + r]] = c[r];
#pragma unroll
          for (int i = 0; i < TB; ++i) o[w[kHdrG + i]] = c[R + i];
          o[K] = c[K];
        }
      } else {
        const int wave = tid >> 6;
        if ((tid & 63) == 0) {
#pragma unroll
          for (int e = 0; e <= K; ++e) red[wave * (K + 1) + e] = c[e];
        }
        __syncthreads();
        const int wps = 1 << (seglog - 6);  // waves per sample
        const int nsamp = (T / 64) / wps;
        if (tid < nsamp * (K + 1)) {
          const int sm = tid / (K + 1), e = tid % (K + 1);
          float x = 0.f;
          for (int q = 0; q < wps; ++q) x += red[(sm * wps + q) * (K + 1) + e];
          int dst;
          if (e < R)
            dst = w[kHdrP + e];
          else if (e < K)
            dst = w[kHdrG + e - R];
          else
            dst = K;
          mp[sm * (K + 1) + dst] = x;
        }
      }
    }

#pragma unroll
    for (int j = 0; j < NR; ++j) tile[addr(j)] = v[j];
    __syncthreads();
    w = op;
  }

  if (A->store) {
#pragma unroll
    for (int it = 0; it < (NR >> 1); ++it) {
      int64_t g_it = 0;
#pragma unroll
      for (int i = 0; i < R - 1; ++i)
        if ((it >> i) & 1) g_it |= int64_t(1) << A->lpos[TB + 1 + i];
      const int64_t gi = base | g_tid | g_it;
      const int sl = swz(2 * (it * T + tid));
      float4 val = *reinterpret_cast<const float4 *>(&tile[sl & ~1]);
      if (sl & 1) val = make_float4(val.z, val.w, val.x, val.y);
      if (gi < total) store_g4(A->state, gi >> 1, val);
    }
  }
}

// Waves per SIMD the pass kernels are register-limited to: two for 32
// amplitudes per thread (the LDS tile allows about two and a half), three below;
// one for the multi-sample form with 32 amplitudes per thread, which also
// carries the vector-ALU path for per-sample unitaries.
template <int R, bool MULTI = false>
constexpr int kPassWaves = R >= 5 ? (MULTI ? 1 : 2) : 3;

// One pass over `ntiles` tiles.  Each workgroup walks tiles t, t + G, ...
// (G = gridDim.x, sized to the resident workgroups) and keeps the next tile's
// HBM loads in flight while it computes the current one, so the memory
// latency of a tile overlaps the arithmetic of the previous one.
template <int R, int TB, bool MULTI>
__global__ void __launch_bounds__(1 << TB) __attribute__((amdgpu_waves_per_eu(kPassWaves<R, MULTI>)))
k_pass(PassArgs A, int64_t ntiles) {
  __shared__ __align__(16) float2 tile[1 << (R + TB)];
  __shared__ float red[((1 << TB) / 64) * (R + TB + 1)];
  constexpr int K = R + TB;
  const int tid = threadIdx.x;
  const int64_t g_tid = thread_offset<TB>(&A, tid);
  float4 pre[1 << (R - 1)];
  int64_t t = blockIdx.x;
  if (A.src && t < ntiles) fetch_tile<R, TB>(&A, tile_base(&A, t, K, MULTI), g_tid, pre);
  for (; t < ntiles; t += gridDim.x) {
    const int64_t base = tile_base(&A, t, K, MULTI);
    if (A.src) {
      stage_tile<R, TB>(tile, pre, tid);
      __syncthreads();
      const int64_t tn = t + gridDim.x;
      if (tn < ntiles) fetch_tile<R, TB>(&A, tile_base(&A, tn, K, MULTI), g_tid, pre);
    }
    pass_body<R, TB, MULTI>(&A, t, base, g_tid, tile, red);
    __syncthreads();
  }
}

// Work-item form: item i runs tile items[2i+1] of the pass described by
// descs[items[2i]].  Circuits of different widths and programs share a launch;
// the next item's tile is fetched while the current one computes.
template <int R, int TB>
__global__ void __launch_bounds__(1 << TB) __attribute__((amdgpu_waves_per_eu(kPassWaves<R>)))
k_pass_items(const PassArgs *descs, const int32_t *items, int nitems) {
  __shared__ __align__(16) float2 tile[1 << (R + TB)];
  __shared__ float red[((1 << TB) / 64) * (R + TB + 1)];
  constexpr int K = R + TB;
  const int tid = threadIdx.x;
  // descriptors and item list are read-only: uniform reads on the scalar path
  const kptr<PassArgs> D = to_const(descs);
  const kptr<int32_t> I = to_const(items);
  float4 pre[1 << (R - 1)];
  int i = blockIdx.x;
  if (i < nitems) {
    const kptr<PassArgs> A = D + I[2 * i];
    if (A->src) fetch_tile<R, TB>(A, tile_base(A, I[2 * i + 1], K, A->multi), thread_offset<TB>(A, tid), pre);
  }
  for (; i < nitems; i += gridDim.x) {
    const kptr<PassArgs> A = D + I[2 * i];
    const int64_t t = I[2 * i + 1];
    const int64_t base = tile_base(A, t, K, A->multi);
    const int64_t g_tid = thread_offset<TB>(A, tid);
    if (A->src) {
      stage_tile<R, TB>(tile, pre, tid);
      __syncthreads();
    }
    const int in = i + gridDim.x;
    if (in < nitems) {
      const kptr<PassArgs> An = D + I[2 * in];
      if (An->src)
        fetch_tile<R, TB>(An, tile_base(An, I[2 * in + 1], K, An->multi), thread_offset<TB>(An, tid), pre);
    }
    if (A->multi)
      pass_body<R, TB, true>(A, t, base, g_tid, tile, red);
    else
      pass_body<R, TB, false>(A, t, base, g_tid, tile, red);
    __syncthreads();
  }
}

// P[b, q] from the per-tile partials of the pass that measured.
struct FinishArgs {
  float *out;          // (B, n) float32
  const float *mpart;  // partials of the measuring pass
  int64_t B;
  int n;
  int K;
  int multi;
  int nn;
  int32_t loc_of_pos[QS_MAX_QUBITS];  // local bit index of a global position, -1 if not local
  int32_t nl_of_pos[QS_MAX_QUBITS];   // index into the tile number's bits, -1 if local
};

__device__ __forceinline__ void measure_finish_block(const FinishArgs &F, int64_t b, int q, double *red) {
  const int n = F.n, K = F.K;
  const int pos = n - 1 - q;
  double acc = 0.0;
  if (F.multi) {
    if (threadIdx.x == 0) {
      const int64_t t = b >> (K - n), s = b & ((int64_t(1) << (K - n)) - 1);
      acc = F.mpart[((t << (K - n)) + s) * (K + 1) + pos];
    }
  } else {
    const int64_t ntl = int64_t(1) << F.nn;
    const int lp = F.loc_of_pos[pos], np = F.nl_of_pos[pos];
    for (int64_t tl = threadIdx.x; tl < ntl; tl += 256) {
      const float *pp = F.mpart + (b * ntl + tl) * (K + 1);
      if (lp >= 0)
        acc += pp[lp];
      else if (!((tl >> np) & 1))
        acc += pp[K];
    }
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) F.out[b * n + q] = (float)(red[0] + red[1] + red[2] + red[3]);
}

__global__ void __launch_bounds__(256) k_measure_finish(FinishArgs F) {
  __shared__ double red[4];
  measure_finish_block(F, blockIdx.x, blockIdx.y, red);
}

// one block per (circuit, qubit); every sample of the circuit in turn
__global__ void __launch_bounds__(256) k_measure_finish_items(const FinishArgs *descs) {
  __shared__ double red[4];
  const FinishArgs &F = descs[blockIdx.x];
  const int q = blockIdx.y;
  if (q >= F.n) return;
  for (int64_t b = 0; b < F.B; ++b) measure_finish_block(F, b, q, red);
}

// readout partials: one block per (sample, qubit), double accumulation over
// the sample's slots (n + 1 floats each, entry n - 1 - q for qubit q)
__global__ void __launch_bounds__(256) k_readout_finish(float *out, const float *mpart, int n, int64_t sps) {
  __shared__ double red[4];
  const int64_t b = blockIdx.x;
  const int q = blockIdx.y;
  const int pos = n - 1 - q;
  double acc = 0.0;
  for (int64_t s = threadIdx.x; s < sps; s += 256) acc += mpart[(b * sps + s) * (n + 1) + pos];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[b * n + q] = (float)(red[0] + red[1] + red[2] + red[3]);
}

// (B, n, 2) product-state factors of an AngleEmbedding (or |0> for unembedded bits)
struct FactorArgs {
  int32_t slot_of_bit[QS_MAX_QUBITS];
  int rot;
};

__global__ void k_product_factors(float *efac, const float *x, int64_t B, int n, int64_t x_stride,
                                  FactorArgs a) {
  const int64_t count = B * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < count;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / n;
    const int pos = (int)(e % n);
    const int j = a.slot_of_bit[pos];
    float2 f0 = make_float2(1.f, 0.f), f1 = make_float2(0.f, 0.f);
    if (j >= 0 && x) {
      float s, c;
      sincosf(x[b * x_stride + j] * 0.5f, &s, &c);
      if (a.rot == QS_RX) {
        f0 = make_float2(c, 0.f);
        f1 = make_float2(0.f, -s);
      } else if (a.rot == QS_RY) {
        f0 = make_float2(c, 0.f);
        f1 = make_float2(s, 0.f);
      } else {
        f0 = make_float2(c, -s);
      }
    }
    float2 *o = reinterpret_cast<float2 *>(efac) + e * 2;
    o[0] = f0;
    o[1] = f1;
  }
}

// Single-qubit unitaries applied to product-state factors: out[b][p] = U f[b][p]
// with U the utab entry a.slot_of_bit[p] (+ b when bit p of per_sample is set);
// positions without an entry are copied.
__global__ void k_factor_apply(float *out, const float *efac, const float *utab, int64_t B, int n,
                               FactorArgs a, uint64_t per_sample) {
  const int64_t count = B * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < count;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / n;
    const int pos = (int)(e % n);
    const float2 *f = reinterpret_cast<const float2 *>(efac) + e * 2;
    float2 g0 = f[0], g1 = f[1];
    const int u = a.slot_of_bit[pos];
    if (u >= 0) {
      const float *m = utab + 8 * ((int64_t)u + (((per_sample >> pos) & 1) ? b : 0));
      const float2 f0 = g0, f1 = g1;
      g0 = make_float2(m[0] * f0.x - m[1] * f0.y + m[2] * f1.x - m[3] * f1.y,
                       m[0] * f0.y + m[1] * f0.x + m[2] * f1.y + m[3] * f1.x);
      g1 = make_float2(m[4] * f0.x - m[5] * f0.y + m[6] * f1.x - m[7] * f1.y,
                       m[4] * f0.y + m[5] * f0.x + m[6] * f1.y + m[7] * f1.x);
    }
    float2 *o = reinterpret_cast<float2 *>(out) + e * 2;
    o[0] = g0;
    o[1] = g1;
  }
}

// Factors of many product states at once: entry e (one bit position of one
// sample of one circuit) reads x[entries[2e]] (or is |0> when that index is
// -1) with rotation entries[2e + 1].
__global__ void k_factor_entries(float *efac, const float *x, const int32_t *entries, int64_t count) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < count;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int xi = entries[2 * e], rot = entries[2 * e + 1];
    float2 f0 = make_float2(1.f, 0.f), f1 = make_float2(0.f, 0.f);
    if (xi >= 0) {
      float s, c;
      sincosf(x[xi] * 0.5f, &s, &c);
      if (rot == QS_RX) {
        f0 = make_float2(c, 0.f);
        f1 = make_float2(0.f, -s);
      } else if (rot == QS_RY) {
        f0 = make_float2(c, 0.f);
        f1 = make_float2(s, 0.f);
      } else {
        f0 = make_float2(c, -s);
      }
    }
    reinterpret_cast<float4 *>(efac)[e] = make_float4(f0.x, f0.y, f1.x, f1.y);
  }
}

// ---------------------------------------------------------------------------
// Table composition: fused per-bit unitaries and their gradient generators.
// Recipe r (at words + off[r]): u_index, span, nconst, q_index, then per
// constituent (kind, cs_index, cs_per_sample) in application order.  For every
// sample variant b < span (the circuit's batch for a per-sample entry, else 1):
//   utab[u_index + b]        = R_m ... R_2 R_1
//   qtab[q_index + k*span + b] = V_k P_k V_k^dagger,  V_k = R_m ... R_{k+1}
// (the generator that turns the cross density at the unitary's output into
// Im<lam|P_k|psi> at constituent k's output).  kind >= 8 is a gradient probe:
// identity rotation with Pauli kind - 8.
// ---------------------------------------------------------------------------
struct C2 {
  double re, im;
};
__device__ __forceinline__ C2 cm(C2 a, C2 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ C2 ca(C2 a, C2 b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ C2 cc(C2 a) { return {a.re, -a.im}; }
struct M2 {
  C2 a, b, c, d;  // [[a, b], [c, d]]
};
__device__ __forceinline__ M2 mm(const M2 &x, const M2 &y) {
  return {ca(cm(x.a, y.a), cm(x.b, y.c)), ca(cm(x.a, y.b), cm(x.b, y.d)),
          ca(cm(x.c, y.a), cm(x.d, y.c)), ca(cm(x.c, y.b), cm(x.d, y.d))};
}
__device__ __forceinline__ M2 mdag(const M2 &x) { return {cc(x.a), cc(x.c), cc(x.b), cc(x.d)}; }

__device__ __forceinline__ M2 rot_matrix(int kind, float c, float s) {
  const C2 z = {0, 0};
  if (kind == QS_RX) return {{c, 0}, {0, -s}, {0, -s}, {c, 0}};
  if (kind == QS_RY) return {{c, 0}, {-s, 0}, {s, 0}, {c, 0}};
  if (kind == QS_RZ) return {{c, -s}, z, z, {c, s}};
  return {{1, 0}, z, z, {1, 0}};  // gradient probe
}
__device__ __forceinline__ M2 pauli(int p) {
  const C2 z = {0, 0};
  if (p == QS_RX) return {z, {1, 0}, {1, 0}, z};
  if (p == QS_RY) return {z, {0, -1}, {0, 1}, z};
  return {{1, 0}, z, z, {-1, 0}};
}
__device__ __forceinline__ void store_m2(float *o, const M2 &x) {
  o[0] = (float)x.a.re; o[1] = (float)x.a.im; o[2] = (float)x.b.re; o[3] = (float)x.b.im;
  o[4] = (float)x.c.re; o[5] = (float)x.c.im; o[6] = (float)x.d.re; o[7] = (float)x.d.im;
}

__device__ __forceinline__ M2 recipe_rot(const int32_t *w, int k, int64_t b, const float2 *csv, int *p) {
  const int kind = w[4 + 3 * k], ci = w[5 + 3 * k], cps = w[6 + 3 * k];
  *p = kind < 8 ? kind : kind - 8;
  if (kind >= 8) return rot_matrix(-1, 1.f, 0.f);
  const float2 v = csv[ci + (cps ? b : 0)];
  return rot_matrix(kind, v.x, v.y);
}

// Per-lane matrix-core operands of fused unitaries: for v_mfma_f32_4x4x1f32,
// lane row r = lane & 3 of a 4-lane block supplies A[r][k] of the real 4x4 form
// of U (or, with dagger, of U^dagger, its transpose).  Entry e (16 floats,
// ltab[16 e + 4 r + k]) is made from utab entry idx[e] (e when idx is NULL).
__global__ void k_lane_coeffs(float *ltab, const float *utab, const int32_t *idx, int64_t count, int dagger) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < count * 4;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t u = e >> 2;
    const int r = (int)(e & 3);
    const float *m = utab + 8 * (idx ? (int64_t)idx[u] : u);
    const float a = m[0], b = m[1], c = m[2], d = m[3], f = m[5], g = m[6], h = m[7], ee = m[4];
    float4 o;
    if (!dagger) {
      if (r == 0) o = make_float4(a, -b, c, -d);
      else if (r == 1) o = make_float4(b, a, d, c);
      else if (r == 2) o = make_float4(ee, -f, g, -h);
      else o = make_float4(f, ee, h, g);
    } else {
      if (r == 0) o = make_float4(a, b, ee, f);
      else if (r == 1) o = make_float4(-b, a, -f, ee);
      else if (r == 2) o = make_float4(c, d, g, h);
      else o = make_float4(-d, c, -h, g);
    }
    reinterpret_cast<float4 *>(ltab)[e] = o;
  }
}

__global__ void __launch_bounds__(64)
k_compose(float *utab, float *qtab, const int32_t *words, const int32_t *off, int nrec,
          const float *cs) {
  const float2 *csv = reinterpret_cast<const float2 *>(cs);
  for (int r = blockIdx.x; r < nrec; r += gridDim.x) {
    const int32_t *w = words + off[r];
    const int64_t uidx = w[0];
    const int64_t span = w[1];
    const int m = w[2];
    const int64_t qidx = w[3];
    for (int64_t b = threadIdx.x; b < span; b += blockDim.x) {
      int p;
      M2 U = {{1, 0}, {0, 0}, {0, 0}, {1, 0}};
      for (int k = 0; k < m; ++k) U = mm(recipe_rot(w, k, b, csv, &p), U);
      store_m2(utab + 8 * (uidx + b), U);
      if (!qtab) continue;
      M2 V = {{1, 0}, {0, 0}, {0, 0}, {1, 0}};
      for (int k = m - 1; k >= 0; --k) {
        const M2 Rk = recipe_rot(w, k, b, csv, &p);
        const M2 Q = mm(mm(V, pauli(p)), mdag(V));
        store_m2(qtab + 8 * (qidx + k * span + b), Q);
        V = mm(V, Rk);
      }
    }
  }
}

template <int R, int TB>
static int launch_pass(const PassArgs &A, int64_t ntiles, hipStream_t st) {
  if (A.multi) {
    const unsigned g = resident_grid(k_pass<R, TB, true>, 1 << TB, ntiles);
    hipLaunchKernelGGL((k_pass<R, TB, true>), dim3(g), dim3(1 << TB), 0, st, A, ntiles);
  } else {
    const unsigned g = resident_grid(k_pass<R, TB, false>, 1 << TB, ntiles);
    hipLaunchKernelGGL((k_pass<R, TB, false>), dim3(g), dim3(1 << TB), 0, st, A, ntiles);
  }
  return launch_status();
}

}  // namespace qs

using namespace qs;

static int64_t fused_tiles(int64_t B, int n, int K) {
  if (n >= K) return B << (n - K);
  const int64_t spt = int64_t(1) << (K - n);
  return (B + spt - 1) / spt;
}

extern "C" {

int qs_fused_supported(int R, int TB) {
  return (TB == 8 && R >= 1 && R <= 5) || (TB == 7 && (R == 5 || R == 4)) || (TB == 6 && R == 5);
}

int qs_layout_supported(int R, int TB) {
  return R >= 1 && R <= 5 && TB >= 6 && TB <= 10 && R + TB <= QS_MAX_LOCAL_BITS;
}

int64_t qs_fused_measure_work_bytes(int64_t B, int n, int R, int TB) {
  if (B < 1 || n < 1 || n > QS_MAX_QUBITS || !qs_layout_supported(R, TB)) return -1;
  const int K = R + TB;
  const int64_t spt = n >= K ? 1 : (int64_t(1) << (K - n));
  return fused_tiles(B, n, K) * spt * (K + 1) * (int64_t)sizeof(float);
}

}  // extern "C"

// Fills a PassArgs from the public arguments (shared by the direct launch and
// the work-item descriptors).  Returns QS_OK or QS_EINVAL.
static int fill_pass_args(PassArgs &A, float *state, const float *src, int64_t B, int n, int R, int TB,
                          const int32_t *lpos, const int32_t *words, int nsub, const float *cs,
                          const float *utab, const float *efac, float *mpart, int store) {
  if (B < 1 || n < 1 || n > QS_MAX_QUBITS || !state || !lpos || !words || nsub < 1 ||
      !qs_fused_supported(R, TB) || R > n || (!src && !efac))
    return QS_EINVAL;
  const int K = R + TB;
  A.state = state;
  A.src = src;
  A.cs = reinterpret_cast<const int32_t *>(cs);
  A.utab = reinterpret_cast<const int32_t *>(utab);
  A.efac = efac;
  A.mpart = mpart;
  A.words = words;
  A.B = B;
  A.n = n;
  A.nsub = nsub;
  A.store = store;
  A.multi = n < K;
  for (int j = 0; j < QS_MAX_LOCAL_BITS; ++j) A.lpos[j] = 0;
  for (int j = 0; j < QS_MAX_QUBITS; ++j) A.npos[j] = 0;
  if (A.multi) {
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
    if (A.lpos[0] != 0) return QS_EINVAL;  // amplitude pairs must be contiguous
    int c = 0;
    for (int p = 0; p < n; ++p)
      if (!((seen >> p) & 1)) A.npos[c++] = p;
    A.nn = c;
  }
  return QS_OK;
}

extern "C" {

int qs_fused_pass(float *state, const float *src, int64_t B, int n, int R, int TB,
                  const int32_t *lpos, const int32_t *words, int nsub, const float *cs,
                  const float *utab, const float *efac, float *mpart, int store, void *stream) {
  PassArgs A;
  int st0 = fill_pass_args(A, state, src, B, n, R, TB, lpos, words, nsub, cs, utab, efac, mpart, store);
  if (st0) return st0;
  const int K = R + TB;
  const int64_t ntiles = fused_tiles(B, n, K);
  if (ntiles > 0x7fffffff) return QS_EUNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  if (TB == 8) {
    switch (R) {
      case 1: return launch_pass<1, 8>(A, ntiles, st);
      case 2: return launch_pass<2, 8>(A, ntiles, st);
      case 3: return launch_pass<3, 8>(A, ntiles, st);
      case 4: return launch_pass<4, 8>(A, ntiles, st);
      case 5: return launch_pass<5, 8>(A, ntiles, st);
    }
  }
  if (TB == 7 && R == 5) return launch_pass<5, 7>(A, ntiles, st);
  if (TB == 7 && R == 4) return launch_pass<4, 7>(A, ntiles, st);
  if (TB == 6 && R == 5) return launch_pass<5, 6>(A, ntiles, st);
  return QS_EUNSUPPORTED;
}

int qs_compose_tables(float *utab, float *qtab, const int32_t *recipes, const int32_t *offsets,
                      int nrec, const float *cs, void *stream) {
  if (nrec < 0 || (nrec > 0 && (!utab || !recipes || !offsets || !cs)))
    return QS_EINVAL;
  if (nrec == 0) return QS_OK;
  const unsigned g = (unsigned)(nrec < 65536 ? nrec : 65536);
  hipLaunchKernelGGL(k_compose, dim3(g), dim3(64), 0, (hipStream_t)stream, utab, qtab, recipes, offsets,
                     nrec, cs);
  return launch_status();
}

int qs_lane_coeffs(float *ltab, const float *utab, const int32_t *idx, int64_t count, int dagger,
                   void *stream) {
  if (count < 0 || (count > 0 && (!ltab || !utab))) return QS_EINVAL;
  if (count == 0) return QS_OK;
  int64_t blocks = (count * 4 + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_lane_coeffs, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, ltab, utab, idx,
                     count, dagger);
  return launch_status();
}

int qs_factor_entries(float *efac, const float *x, const int32_t *entries, int64_t count,
                      void *stream) {
  if (count < 0 || (count > 0 && (!efac || !entries))) return QS_EINVAL;
  if (count == 0) return QS_OK;
  int64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_factor_entries, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     efac, x, entries, count);
  return launch_status();
}

int64_t qs_pass_desc_bytes(void) { return (int64_t)sizeof(PassArgs); }

int qs_pass_desc(void *desc, int64_t *ntiles, float *state, const float *src, int64_t B, int n,
                 int R, int TB, const int32_t *lpos, const int32_t *words, int nsub, const float *cs,
                 const float *utab, const float *efac, float *mpart, int store) {
  if (!desc || !ntiles) return QS_EINVAL;
  PassArgs A;
  int st0 = fill_pass_args(A, state, src, B, n, R, TB, lpos, words, nsub, cs, utab, efac, mpart, store);
  if (st0) return st0;
  *ntiles = fused_tiles(B, n, R + TB);
  memcpy(desc, &A, sizeof(PassArgs));
  return QS_OK;
}

int qs_items_pass(const void *descs, const int32_t *items, int nitems, int R, int TB, int grid,
                  void *stream) {
  if (!descs || !items || nitems < 0 || grid < 1) return QS_EINVAL;
  if (nitems == 0) return QS_OK;
  hipStream_t st = (hipStream_t)stream;
  const PassArgs *d = (const PassArgs *)descs;
  const unsigned g = (unsigned)(grid < nitems ? grid : nitems);
  if (R == 4 && TB == 7) {
    hipLaunchKernelGGL((k_pass_items<4, 7>), dim3(g), dim3(128), 0, st, d, items, nitems);
    return launch_status();
  }
  if (R == 4 && TB == 8) {
    hipLaunchKernelGGL((k_pass_items<4, 8>), dim3(g), dim3(256), 0, st, d, items, nitems);
    return launch_status();
  }
  if (R == 5 && TB == 6) {
    hipLaunchKernelGGL((k_pass_items<5, 6>), dim3(g), dim3(64), 0, st, d, items, nitems);
    return launch_status();
  }
  if (R == 5 && TB == 8) {
    hipLaunchKernelGGL((k_pass_items<5, 8>), dim3(g), dim3(256), 0, st, d, items, nitems);
    return launch_status();
  }
  return QS_EUNSUPPORTED;
}

}  // extern "C"

static int fill_finish_args(FinishArgs &F, float *out, const float *mpart, int64_t B, int n, int R,
                            int TB, const int32_t *lpos) {
  if (B < 1 || n < 1 || n > QS_MAX_QUBITS || !out || !mpart || !lpos || !qs_layout_supported(R, TB))
    return QS_EINVAL;
  const int K = R + TB;
  F.out = out;
  F.mpart = mpart;
  F.B = B;
  F.n = n;
  F.K = K;
  F.multi = n < K;
  for (int p = 0; p < QS_MAX_QUBITS; ++p) F.loc_of_pos[p] = F.nl_of_pos[p] = -1;
  int nn = 0;
  if (F.multi) {
    for (int p = 0; p < n; ++p) F.loc_of_pos[p] = p;
  } else {
    for (int j = 0; j < K; ++j) {
      if (lpos[j] < 0 || lpos[j] >= n) return QS_EINVAL;
      F.loc_of_pos[lpos[j]] = j;
    }
    for (int p = 0; p < n; ++p)
      if (F.loc_of_pos[p] < 0) F.nl_of_pos[p] = nn++;
  }
  F.nn = nn;
  return QS_OK;
}

extern "C" {

int qs_fused_measure_finish(float *out, const float *mpart, int64_t B, int n, int R, int TB,
                            const int32_t *lpos, void *stream) {
  FinishArgs F;
  int st0 = fill_finish_args(F, out, mpart, B, n, R, TB, lpos);
  if (st0) return st0;
  if (B > 0x7fffffff) return QS_EUNSUPPORTED;
  hipLaunchKernelGGL(k_measure_finish, dim3((unsigned)B, (unsigned)n), dim3(256), 0,
                     (hipStream_t)stream, F);
  return launch_status();
}

int64_t qs_readout_work_bytes(int64_t B, int n, int R, int TB) {
  if (B < 1 || n < 1 || n > QS_MAX_QUBITS || !qs_layout_supported(R, TB)) return -1;
  const int K = R + TB;
  const int64_t spt = n >= K ? 1 : (int64_t(1) << (K - n));
  return fused_tiles(B, n, K) * spt * (n + 1) * (int64_t)sizeof(float);
}

int qs_readout_finish(float *out, const float *mpart, int64_t B, int n, int R, int TB, void *stream) {
  if (B < 1 || n < 1 || n > QS_MAX_QUBITS || !out || !mpart || !qs_layout_supported(R, TB)) return QS_EINVAL;
  if (B > 0x7fffffff) return QS_EUNSUPPORTED;
  const int K = R + TB;
  const int64_t sps = n >= K ? (int64_t(1) << (n - K)) : 1;
  hipLaunchKernelGGL(k_readout_finish, dim3((unsigned)B, (unsigned)n), dim3(256), 0, (hipStream_t)stream, out,
                     mpart, n, sps);
  return launch_status();
}

int64_t qs_measure_desc_bytes(void) { return (int64_t)sizeof(FinishArgs); }

int qs_measure_desc(void *desc, float *out, const float *mpart, int64_t B, int n, int R, int TB,
                    const int32_t *lpos) {
  if (!desc) return QS_EINVAL;
  FinishArgs F;
  int st0 = fill_finish_args(F, out, mpart, B, n, R, TB, lpos);
  if (st0) return st0;
  memcpy(desc, &F, sizeof(FinishArgs));
  return QS_OK;
}

int qs_items_measure_finish(const void *descs, int ndesc, int max_n, void *stream) {
  if (!descs || ndesc < 0 || max_n < 1 || max_n > QS_MAX_QUBITS) return QS_EINVAL;
  if (ndesc == 0) return QS_OK;
  hipLaunchKernelGGL(k_measure_finish_items, dim3((unsigned)ndesc, (unsigned)max_n), dim3(256), 0,
                     (hipStream_t)stream, (const FinishArgs *)descs);
  return launch_status();
}

int qs_product_factors(float *efac, const float *x, int64_t B, int n, const int32_t *bits, int d,
                       int rot, int64_t x_stride, void *stream) {
  if (B < 1 || n < 1 || n > QS_MAX_QUBITS || !efac || d < 0 || d > n || rot < 0 || rot > 2 ||
      (d > 0 && (!x || !bits)))
    return QS_EINVAL;
  FactorArgs a;
  for (int p = 0; p < QS_MAX_QUBITS; ++p) a.slot_of_bit[p] = -1;
  for (int j = 0; j < d; ++j) {
    if (bits[j] < 0 || bits[j] >= n || a.slot_of_bit[bits[j]] >= 0) return QS_EINVAL;
    a.slot_of_bit[bits[j]] = j;
  }
  a.rot = rot;
  const int64_t count = B * n;
  int64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_product_factors, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     efac, d > 0 ? x : nullptr, B, n, x_stride, a);
  return launch_status();
}

int qs_factor_apply(float *out, const float *efac, const float *utab, int64_t nu, const int32_t *unit,
                    const int32_t *span, int64_t B, int n, void *stream) {
  if (B < 1 || n < 1 || n > QS_MAX_QUBITS || !out || !efac || !unit || !span || nu < 0) return QS_EINVAL;
  FactorArgs a;
  uint64_t per_sample = 0;
  for (int p = 0; p < n; ++p) {
    const int32_t u = unit[p];
    a.slot_of_bit[p] = -1;
    if (u < 0) continue;
    if (!utab || (span[p] != 1 && span[p] != B) || (int64_t)u + span[p] > nu) return QS_EINVAL;
    a.slot_of_bit[p] = u;
    if (span[p] == B && B > 1) per_sample |= uint64_t(1) << p;
  }
  a.rot = 0;
  const int64_t count = B * n;
  int64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_factor_apply, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, out, efac,
                     utab, B, n, a, per_sample);
  return launch_status();
}

}  // extern "C"
