// Plain per-gate path for gfx950: one HBM pass per gate.
//
// These kernels are what a single qandle operator module costs when it runs on
// its own (BuiltRX.forward, BuiltCNOT.forward, ...).  The fused program path
// (qs_fused.hip) is what Circuit.forward uses; this file is also its
// correctness and speed comparator.
#include <mutex>

#include "qs_common.h"

namespace qs {
int g_last_hip_error = 0;

constexpr int kThreads = 256;
constexpr int kMaxBits = 64;

static inline unsigned grid_for(int64_t work_items, int per_thread = 1) {
  int64_t blocks = (work_items + (int64_t)kThreads * per_thread - 1) / ((int64_t)kThreads * per_thread);
  if (blocks < 1) blocks = 1;
  if (blocks > (1 << 20)) blocks = 1 << 20;  // grid-stride loops cover the rest
  return (unsigned)blocks;
}

// ----------------------------------------------------------------- coeffs
__global__ void k_half_angle(float *cs, const float *phi, int64_t count) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count;
       k += (int64_t)gridDim.x * blockDim.x) {
    float s, c;
    sincosf(phi[k] * 0.5f, &s, &c);
    cs[2 * k] = c;
    cs[2 * k + 1] = s;
  }
}

// ------------------------------------------------------------------ embed
struct EmbedArgs {
  int32_t slot_of_bit[kMaxBits];  // embedded column for each bit position, -1 if none
  int d;
  int rot;
};

// Factors (f(0), f(1)) of one embedded qubit: column 0 of R_rot(x).
__device__ __forceinline__ void embed_factor(int rot, float x, float2 &f0, float2 &f1) {
  float s, c;
  sincosf(x * 0.5f, &s, &c);
  if (rot == QS_RX) {
    f0 = make_float2(c, 0.f);
    f1 = make_float2(0.f, -s);
  } else if (rot == QS_RY) {
    f0 = make_float2(c, 0.f);
    f1 = make_float2(s, 0.f);
  } else {
    f0 = make_float2(c, -s);
    f1 = make_float2(0.f, 0.f);
  }
}

// Each thread writes 2^L consecutive amplitudes (L = min(3, n)).  A block
// covers kThreads << L amplitudes; the factor table of every sample the block
// touches is staged in LDS once.
constexpr int kEmbedMaxSamplesPerBlock = kThreads * 8;  // n = 0 worst case
__global__ void __launch_bounds__(kThreads)
k_embed(float *out, const float *x, int64_t B, int n, EmbedArgs a, int64_t x_stride) {
  extern __shared__ float2 tab[];  // [spb][n][2]
  const int L = n < 3 ? n : 3;
  const int64_t amps_per_block = (int64_t)kThreads << L;
  const int64_t total = B << n;
  const int64_t first = blockIdx.x * amps_per_block;
  const int64_t s0 = first >> n;
  int64_t s_last = (first + amps_per_block - 1) >> n;
  if (s_last >= B) s_last = B - 1;
  const int spb = (int)(s_last - s0 + 1);
  for (int e = threadIdx.x; e < spb * n; e += kThreads) {
    int sl = e / n, p = e % n;
    int j = a.slot_of_bit[p];
    float2 f0 = make_float2(1.f, 0.f), f1 = make_float2(0.f, 0.f);
    if (j >= 0) embed_factor(a.rot, x[(s0 + sl) * x_stride + j], f0, f1);
    tab[(sl * n + p) * 2 + 0] = f0;
    tab[(sl * n + p) * 2 + 1] = f1;
  }
  __syncthreads();
  const int64_t i_first = first + ((int64_t)threadIdx.x << L);
  if (i_first >= total) return;
  const int64_t b = i_first >> n;
  const int64_t local = i_first & ((int64_t(1) << n) - 1);
  const float2 *t = tab + (int)(b - s0) * n * 2;
  float2 pre = make_float2(1.f, 0.f);
  for (int p = L; p < n; ++p) pre = cmul(pre, t[p * 2 + ((local >> p) & 1)]);
  float2 v[8];
  v[0] = pre;
  int cnt = 1;
  for (int p = L - 1; p >= 0; --p) {  // doubling: bit p becomes the next-lower index bit
    for (int k = cnt - 1; k >= 0; --k) {
      v[2 * k + 1] = cmul(v[k], t[p * 2 + 1]);
      v[2 * k] = cmul(v[k], t[p * 2 + 0]);
    }
    cnt *= 2;
  }
  for (int k = 0; k < cnt; ++k) st2(out, i_first + k, v[k]);
}

// -------------------------------------------------------------------- rot
template <int KIND>
__global__ void __launch_bounds__(kThreads)
k_rot(float *out, const float *in, int64_t npairs, int n, int bit, const float *cs,
      int64_t cs_stride, float sgn) {
  const int64_t m = int64_t(1) << bit;
  for (int64_t p = blockIdx.x * (int64_t)kThreads + threadIdx.x; p < npairs;
       p += (int64_t)gridDim.x * kThreads) {
    int64_t i0 = insert_zero(p, bit), i1 = i0 | m;
    int64_t b = i0 >> n;
    float2 c_s = ld2(cs, b * cs_stride);
    float2 a0 = ld2(in, i0), a1 = ld2(in, i1);
    rot_pair<KIND>(c_s.x, sgn * c_s.y, a0, a1);
    st2(out, i0, a0);
    st2(out, i1, a1);
  }
}

// bit 0: the pair is one 16-byte vector
template <int KIND>
__global__ void __launch_bounds__(kThreads)
k_rot_bit0(float *out, const float *in, int64_t npairs, int n, const float *cs,
           int64_t cs_stride, float sgn) {
  for (int64_t p = blockIdx.x * (int64_t)kThreads + threadIdx.x; p < npairs;
       p += (int64_t)gridDim.x * kThreads) {
    int64_t b = (2 * p) >> n;
    float2 c_s = ld2(cs, b * cs_stride);
    float4 v = reinterpret_cast<const float4 *>(in)[p];
    float2 a0 = make_float2(v.x, v.y), a1 = make_float2(v.z, v.w);
    rot_pair<KIND>(c_s.x, sgn * c_s.y, a0, a1);
    reinterpret_cast<float4 *>(out)[p] = make_float4(a0.x, a0.y, a1.x, a1.y);
  }
}

// ------------------------------------------------------------------- cnot
__global__ void __launch_bounds__(kThreads)
k_cnot_gather(float *out, const float *in, int64_t total, int64_t cm, int64_t tm) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kThreads) {
    int64_t src = (i & cm) ? (i ^ tm) : i;
    st2(out, i, ld2(in, src));
  }
}

__global__ void __launch_bounds__(kThreads)
k_cnot_swap(float *psi, int64_t quarter, int lo, int hi, int64_t cm, int64_t tm) {
  for (int64_t p = blockIdx.x * (int64_t)kThreads + threadIdx.x; p < quarter;
       p += (int64_t)gridDim.x * kThreads) {
    int64_t i = insert_zero(insert_zero(p, lo), hi) | cm;
    float2 a = ld2(psi, i), b = ld2(psi, i | tm);
    st2(psi, i, b);
    st2(psi, i | tm, a);
  }
}

// ---------------------------------------------------------- reductions
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// block sum in a fixed order (deterministic); all threads get the result
__device__ float block_sum(float v, float *red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < kThreads / 64; ++k) s += red[k];
  return s;
}

__device__ double block_sum_d(double v, double *red) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kThreads / 64; ++k) s += red[k];
  return s;
}

constexpr int kChunkBits = 12;  // amplitudes reduced per block in stage 1
static inline int chunk_bits(int n) { return n < kChunkBits ? n : kChunkBits; }

// Stage 1 of MeasureProbability: chunk ch (2^c amplitudes of one sample)
// -> part[ch][pos] = sum over the chunk with bit pos = 0 (pos < c), part[ch][c] = total.
__global__ void __launch_bounds__(kThreads)
k_measure_stage1(float *part, const float *psi, int c) {
  __shared__ float red[kThreads / 64];
  const int64_t ch = blockIdx.x;
  const int64_t base = ch << c;
  const int cnt = 1 << c;
  float S[kChunkBits];
#pragma unroll
  for (int p = 0; p < kChunkBits; ++p) S[p] = 0.f;
  float T = 0.f;
  for (int j = threadIdx.x; j < cnt; j += kThreads) {
    float2 a = ld2(psi, base + j);
    float pr = a.x * a.x + a.y * a.y;
    T += pr;
#pragma unroll
    for (int p = 0; p < kChunkBits; ++p)
      if (p < c && !((j >> p) & 1)) S[p] += pr;
  }
  float *out = part + ch * (c + 1);
#pragma unroll
  for (int p = 0; p < kChunkBits; ++p) {
    if (p < c) {
      float s = block_sum(S[p], red);
      if (threadIdx.x == 0) out[p] = s;
    }
  }
  float t = block_sum(T, red);
  if (threadIdx.x == 0) out[c] = t;
}

// Stage 2: block (b, q) sums the chunks of sample b in double, fixed order.
__global__ void __launch_bounds__(kThreads)
k_measure_stage2(float *out, const float *part, int n, int c) {
  __shared__ double red[kThreads / 64];
  const int64_t b = blockIdx.x;
  const int q = blockIdx.y;
  const int pos = n - 1 - q;
  const int64_t nch = int64_t(1) << (n - c);
  double acc = 0.0;
  for (int64_t k = threadIdx.x; k < nch; k += kThreads) {
    const float *pp = part + (b * nch + k) * (c + 1);
    if (pos < c)
      acc += pp[pos];
    else if (!((k >> (pos - c)) & 1))
      acc += pp[c];
  }
  double s = block_sum_d(acc, red);
  if (threadIdx.x == 0) out[b * n + q] = (float)s;
}

// lam = 2 psi * sum_q w[b,q] [bit n-1-q = 0]
__global__ void __launch_bounds__(kThreads)
k_measure_backward(float *lam, const float *psi, const float *w, int64_t w_stride,
                   int64_t total, int n) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kThreads) {
    int64_t b = i >> n;
    const float *wb = w + b * w_stride;
    float acc = 0.f;
    for (int q = 0; q < n; ++q)
      if (!((i >> (n - 1 - q)) & 1)) acc += wb[q];
    float2 a = ld2(psi, i);
    st2(lam, i, make_float2(2.f * acc * a.x, 2.f * acc * a.y));
  }
}

struct PauliArgs {
  int32_t bits[kMaxBits];
  int d;
  int64_t cmask;  // only amplitudes with every bit of cmask set contribute (0 = all)
};

// Stage 1 of Im<lam|P_j|psi>: part[ch][j]
template <int KIND>
__global__ void __launch_bounds__(kThreads)
k_pauli_stage1(float *part, const float *lam, const float *psi, int c, PauliArgs a) {
  __shared__ float red[kThreads / 64];
  const int64_t base = (int64_t)blockIdx.x << c;
  const int cnt = 1 << c;
  for (int j = 0; j < a.d; ++j) {
    const int64_t m = int64_t(1) << a.bits[j];
    float acc = 0.f;
    for (int t = threadIdx.x; t < cnt; t += kThreads) {
      int64_t i = base + t;
      float2 l = ld2(lam, i);
      bool one = (i & m) != 0;
      float2 pp;
      if (KIND == QS_RX) {
        pp = ld2(psi, i ^ m);
      } else if (KIND == QS_RY) {
        float2 q = ld2(psi, i ^ m);  // bit 0: -i q, bit 1: +i q
        pp = one ? make_float2(-q.y, q.x) : make_float2(q.y, -q.x);
      } else {
        float2 q = ld2(psi, i);
        pp = one ? make_float2(-q.x, -q.y) : q;
      }
      if ((i & a.cmask) == a.cmask) acc += l.x * pp.y - l.y * pp.x;  // Im(conj(l) pp)
    }
    float s = block_sum(acc, red);
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * a.d + j] = s;
  }
}

__global__ void __launch_bounds__(kThreads)
k_pauli_stage2(float *out, const float *part, int n, int c, int d) {
  __shared__ double red[kThreads / 64];
  const int64_t b = blockIdx.x;
  const int j = blockIdx.y;
  const int64_t nch = int64_t(1) << (n - c);
  double acc = 0.0;
  for (int64_t k = threadIdx.x; k < nch; k += kThreads) acc += part[(b * nch + k) * d + j];
  double s = block_sum_d(acc, red);
  if (threadIdx.x == 0) out[b * d + j] = (float)s;
}

}  // namespace qs

using namespace qs;

static bool bad_shape(int64_t B, int n) { return B < 1 || n < 1 || n > 40; }

extern "C" {

int qs_abi_version(void) { return 1; }
int qs_last_hip_error(void) { return qs::g_last_hip_error; }

const char *qs_status_string(int status) {
  switch (status) {
    case QS_OK: return "ok";
    case QS_EINVAL: return "invalid argument";
    case QS_EHIP: return "HIP runtime error";
    case QS_EWORKSPACE: return "workspace too small";
    case QS_EUNSUPPORTED: return "unsupported request";
    default: return "unknown status";
  }
}

int qs_half_angle_coeffs(float *cs, const float *phi, int64_t count, void *stream) {
  if (count < 0 || (count > 0 && (!cs || !phi))) return QS_EINVAL;
  if (count == 0) return QS_OK;
  hipLaunchKernelGGL(k_half_angle, dim3(grid_for(count)), dim3(kThreads), 0, (hipStream_t)stream,
                     cs, phi, count);
  return launch_status();
}

int qs_embed_angle(float *out, const float *x, int64_t B, int n, const int32_t *bits, int d,
                   int rot, int64_t x_stride, void *stream) {
  if (bad_shape(B, n) || !out || !x || !bits || d < 1 || d > n || rot < 0 || rot > 2 ||
      x_stride < 0)
    return QS_EINVAL;
  EmbedArgs a;
  for (int p = 0; p < kMaxBits; ++p) a.slot_of_bit[p] = -1;
  for (int j = 0; j < d; ++j) {
    if (bits[j] < 0 || bits[j] >= n || a.slot_of_bit[bits[j]] >= 0) return QS_EINVAL;
    a.slot_of_bit[bits[j]] = j;
  }
  a.d = d;
  a.rot = rot;
  const int L = n < 3 ? n : 3;
  const int64_t total = B << n;
  const int64_t apb = (int64_t)kThreads << L;
  const int64_t blocks = (total + apb - 1) / apb;
  if (blocks > 0x7fffffff) return QS_EUNSUPPORTED;
  int64_t spb = (apb >> n) + 2;  // samples a block can touch
  if (spb > B) spb = B;
  size_t lds = (size_t)spb * n * 2 * sizeof(float2);
  if (lds > 64 * 1024) return QS_EUNSUPPORTED;
  hipLaunchKernelGGL(k_embed, dim3((unsigned)blocks), dim3(kThreads), lds, (hipStream_t)stream,
                     out, x, B, n, a, x_stride);
  return launch_status();
}

int qs_rot(float *out, const float *in, int64_t B, int n, int bit, int kind, const float *cs,
           int64_t cs_stride, int inverse, void *stream) {
  if (bad_shape(B, n) || !out || !in || !cs || bit < 0 || bit >= n || kind < 0 || kind > 2 ||
      cs_stride < 0)
    return QS_EINVAL;
  const int64_t npairs = (B << n) >> 1;
  const float sgn = inverse ? -1.f : 1.f;
  const dim3 g(grid_for(npairs)), blk(kThreads);
  hipStream_t st = (hipStream_t)stream;
  const bool aligned16 = ((uintptr_t)out % 16 == 0) && ((uintptr_t)in % 16 == 0);
  if (bit == 0 && aligned16) {
    if (kind == QS_RX) hipLaunchKernelGGL(k_rot_bit0<QS_RX>, g, blk, 0, st, out, in, npairs, n, cs, cs_stride, sgn);
    if (kind == QS_RY) hipLaunchKernelGGL(k_rot_bit0<QS_RY>, g, blk, 0, st, out, in, npairs, n, cs, cs_stride, sgn);
    if (kind == QS_RZ) hipLaunchKernelGGL(k_rot_bit0<QS_RZ>, g, blk, 0, st, out, in, npairs, n, cs, cs_stride, sgn);
  } else {
    if (kind == QS_RX) hipLaunchKernelGGL(k_rot<QS_RX>, g, blk, 0, st, out, in, npairs, n, bit, cs, cs_stride, sgn);
    if (kind == QS_RY) hipLaunchKernelGGL(k_rot<QS_RY>, g, blk, 0, st, out, in, npairs, n, bit, cs, cs_stride, sgn);
    if (kind == QS_RZ) hipLaunchKernelGGL(k_rot<QS_RZ>, g, blk, 0, st, out, in, npairs, n, bit, cs, cs_stride, sgn);
  }
  return launch_status();
}

int qs_cnot(float *out, const float *in, int64_t B, int n, int cbit, int tbit, void *stream) {
  if (bad_shape(B, n) || !out || !in || cbit < 0 || tbit < 0 || cbit >= n || tbit >= n ||
      cbit == tbit)
    return QS_EINVAL;
  const int64_t total = B << n;
  const int64_t cm = int64_t(1) << cbit, tm = int64_t(1) << tbit;
  hipStream_t st = (hipStream_t)stream;
  if (out == in) {
    if (n < 2) return QS_EINVAL;
    const int lo = cbit < tbit ? cbit : tbit, hi = cbit < tbit ? tbit : cbit;
    const int64_t quarter = total >> 2;
    hipLaunchKernelGGL(k_cnot_swap, dim3(grid_for(quarter)), dim3(kThreads), 0, st, out, quarter,
                       lo, hi, cm, tm);
  } else {
    hipLaunchKernelGGL(k_cnot_gather, dim3(grid_for(total)), dim3(kThreads), 0, st, out, in, total,
                       cm, tm);
  }
  return launch_status();
}

int64_t qs_measure_prob_work_bytes(int64_t B, int n) {
  if (bad_shape(B, n)) return -1;
  const int c = chunk_bits(n);
  return (B << (n - c)) * (c + 1) * (int64_t)sizeof(float);
}

}  // extern "C"

// The two MeasureProbability stages over a partials buffer of
// qs_measure_prob_work_bytes(B, n) bytes (shapes checked by the caller).
static int measure_launch(float *out, const float *psi, int64_t B, int n, float *part, hipStream_t st) {
  const int c = chunk_bits(n);
  const int64_t nch = B << (n - c);
  if (nch > 0x7fffffff || B > 0x7fffffff) return QS_EUNSUPPORTED;
  hipLaunchKernelGGL(k_measure_stage1, dim3((unsigned)nch), dim3(kThreads), 0, st, part, psi, c);
  int s = launch_status();
  if (s) return s;
  hipLaunchKernelGGL(k_measure_stage2, dim3((unsigned)B, (unsigned)n), dim3(kThreads), 0, st, out,
                     (const float *)part, n, c);
  return launch_status();
}

// Stage-1 partials of qs_measure_prob_noalloc.  A file-scope device array: every
// device holds its own copy, at its own address, for the life of the process.
// Host code never passes `scratch` itself to a kernel (on the host that name is
// a stand-in, not a device address); it uses the address looked up per device
// below.  Stage 1 writes every partial that stage 2 reads, so nothing needs
// clearing between calls.
static __device__ __align__(256) float scratch[QS_NOALLOC_SCRATCH_BYTES / sizeof(float)];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64];  // device address of `scratch` per device; null until looked up

// Address of `scratch` on the device of `stream` (the current device for the
// null stream), looked up once per device with that device current.
static int scratch_for(hipStream_t stream, float **out) {
  int dev = 0;
  hipError_t e = stream ? hipStreamGetDevice(stream, &dev) : hipGetDevice(&dev);
  if (e != hipSuccess) {
    g_last_hip_error = (int)e;
    return QS_EHIP;
  }
  if (dev < 0 || dev >= 64) return QS_EUNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    if ((e = hipGetDevice(&cur)) != hipSuccess || (cur != dev && (e = hipSetDevice(dev)) != hipSuccess)) {
      g_last_hip_error = (int)e;
      return QS_EHIP;
    }
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
      g_last_hip_error = (int)e;
      return QS_EHIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<float *>(g_scratch_ptr[dev]);
  return QS_OK;
}

extern "C" {

int qs_measure_prob(float *out, const float *psi, int64_t B, int n, void *work, int64_t work_bytes,
                    void *stream) {
  if (bad_shape(B, n) || !out || !psi || !work) return QS_EINVAL;
  if (work_bytes < qs_measure_prob_work_bytes(B, n)) return QS_EWORKSPACE;
  return measure_launch(out, psi, B, n, (float *)work, (hipStream_t)stream);
}

int qs_measure_prob_noalloc(float *out, const float *psi, int64_t B, int n, void *stream) {
  if (bad_shape(B, n) || !out || !psi) return QS_EINVAL;
  if (qs_measure_prob_work_bytes(B, n) > (int64_t)QS_NOALLOC_SCRATCH_BYTES) return QS_EWORKSPACE;
  float *part = nullptr;
  const int s = scratch_for((hipStream_t)stream, &part);
  if (s) return s;
  return measure_launch(out, psi, B, n, part, (hipStream_t)stream);
}

int qs_measure_prob_backward(float *lam, const float *psi, const float *w, int64_t w_stride,
                             int64_t B, int n, void *stream) {
  if (bad_shape(B, n) || !lam || !psi || !w || w_stride < 0) return QS_EINVAL;
  const int64_t total = B << n;
  hipLaunchKernelGGL(k_measure_backward, dim3(grid_for(total)), dim3(kThreads), 0,
                     (hipStream_t)stream, lam, psi, w, w_stride, total, n);
  return launch_status();
}

int64_t qs_pauli_im_work_bytes(int64_t B, int n, int d) {
  if (bad_shape(B, n) || d < 1) return -1;
  const int c = chunk_bits(n);
  return (B << (n - c)) * (int64_t)d * (int64_t)sizeof(float);
}

int qs_pauli_im_controlled(float *out, const float *lam, const float *psi, int64_t B, int n,
                           const int32_t *bits, int d, int pauli, int64_t cmask, void *work,
                           int64_t work_bytes, void *stream) {
  if (bad_shape(B, n) || !out || !lam || !psi || !bits || !work || d < 1 || d > kMaxBits ||
      pauli < 0 || pauli > 2 || cmask < 0 || (n < 63 && (cmask >> n) != 0))
    return QS_EINVAL;
  if (work_bytes < qs_pauli_im_work_bytes(B, n, d)) return QS_EWORKSPACE;
  PauliArgs a;
  for (int j = 0; j < d; ++j) {
    if (bits[j] < 0 || bits[j] >= n || ((cmask >> bits[j]) & 1)) return QS_EINVAL;
    a.bits[j] = bits[j];
  }
  a.d = d;
  a.cmask = cmask;
  const int c = chunk_bits(n);
  const int64_t nch = B << (n - c);
  if (nch > 0x7fffffff) return QS_EUNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  float *part = (float *)work;
  if (pauli == QS_RX) hipLaunchKernelGGL(k_pauli_stage1<QS_RX>, dim3((unsigned)nch), dim3(kThreads), 0, st, part, lam, psi, c, a);
  if (pauli == QS_RY) hipLaunchKernelGGL(k_pauli_stage1<QS_RY>, dim3((unsigned)nch), dim3(kThreads), 0, st, part, lam, psi, c, a);
  if (pauli == QS_RZ) hipLaunchKernelGGL(k_pauli_stage1<QS_RZ>, dim3((unsigned)nch), dim3(kThreads), 0, st, part, lam, psi, c, a);
  int s = launch_status();
  if (s) return s;
  hipLaunchKernelGGL(k_pauli_stage2, dim3((unsigned)B, (unsigned)d), dim3(kThreads), 0, st, out,
                     (const float *)part, n, c, d);
  return launch_status();
}

int qs_pauli_im(float *out, const float *lam, const float *psi, int64_t B, int n,
                const int32_t *bits, int d, int pauli, void *work, int64_t work_bytes,
                void *stream) {
  return qs_pauli_im_controlled(out, lam, psi, B, n, bits, d, pauli, 0, work, work_bytes, stream);
}

}  // extern "C"
