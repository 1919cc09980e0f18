// Plain per-gate kernels of the remaining operators: arbitrary (optionally
// controlled) one-qubit unitaries, multi-controlled X, Z-type sign flips,
// (controlled) SWAP and Reset.  One HBM pass per gate, like qs_plain.hip.
//
// Every operator of the reference that these replace is a dense row-vector
// product `state @ M` with a 2^n x 2^n matrix (src/qandle/operators.py:125-126,
// 445-447, 488-506, 590-591, 673-674) or a reshaped norm update (:612-623);
// here each is the pair or permutation update that matrix encodes.  Control
// masks are sets of bit positions that must all be 1 for the update to apply
// (0 = uncontrolled).
#include "qs_common.h"
#include "qs_program.h"

namespace qs {

constexpr int kGateThreads = 256;

static inline unsigned gate_grid(int64_t items) {
  int64_t blocks = (items + kGateThreads - 1) / kGateThreads;
  if (blocks < 1) blocks = 1;
  if (blocks > (1 << 20)) blocks = 1 << 20;  // grid-stride loops cover the rest
  return (unsigned)blocks;
}

// column-convention U (row-major complex 2x2: u00, u01, u10, u11) on a pair
__device__ __forceinline__ void apply_u(const float *u, bool dagger, float2 &a0, float2 &a1) {
  float2 u00 = make_float2(u[0], u[1]), u01 = make_float2(u[2], u[3]);
  float2 u10 = make_float2(u[4], u[5]), u11 = make_float2(u[6], u[7]);
  if (dagger) {  // U^dagger = conj(U)^T
    const float2 t = u01;
    u00 = make_float2(u00.x, -u00.y);
    u11 = make_float2(u11.x, -u11.y);
    u01 = make_float2(u10.x, -u10.y);
    u10 = make_float2(t.x, -t.y);
  }
  const float2 n0 = make_float2(u00.x * a0.x - u00.y * a0.y + u01.x * a1.x - u01.y * a1.y,
                                u00.x * a0.y + u00.y * a0.x + u01.x * a1.y + u01.y * a1.x);
  const float2 n1 = make_float2(u10.x * a0.x - u10.y * a0.y + u11.x * a1.x - u11.y * a1.y,
                                u10.x * a0.y + u10.y * a0.x + u11.x * a1.y + u11.y * a1.x);
  a0 = n0;
  a1 = n1;
}

__global__ void __launch_bounds__(kGateThreads)
k_unitary_1q(float *out, const float *in, int64_t npairs, int n, int bit, const float *u, int64_t u_stride,
             int64_t cmask, int dagger) {
  const int64_t m = int64_t(1) << bit;
  for (int64_t p = blockIdx.x * (int64_t)kGateThreads + threadIdx.x; p < npairs;
       p += (int64_t)gridDim.x * kGateThreads) {
    const int64_t i0 = insert_zero(p, bit), i1 = i0 | m;
    float2 a0 = ld2(in, i0), a1 = ld2(in, i1);
    if ((i0 & cmask) == cmask) apply_u(u + 8 * (i0 >> n) * u_stride, dagger != 0, a0, a1);
    st2(out, i0, a0);
    st2(out, i1, a1);
  }
}

// u[8k..8k+8) = R_kind(phi_k) from cs[2k] = cos(phi_k/2), cs[2k+1] = sin(phi_k/2)
__global__ void k_rot_unitaries(float *u, const float *cs, int64_t count, int kind) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count;
       k += (int64_t)gridDim.x * blockDim.x) {
    const float c = cs[2 * k], s = cs[2 * k + 1];
    float *o = u + 8 * k;
    float r[8];
    if (kind == QS_RX) {  // [[c, -is], [-is, c]]
      r[0] = c; r[1] = 0.f; r[2] = 0.f; r[3] = -s; r[4] = 0.f; r[5] = -s; r[6] = c; r[7] = 0.f;
    } else if (kind == QS_RY) {  // [[c, -s], [s, c]]
      r[0] = c; r[1] = 0.f; r[2] = -s; r[3] = 0.f; r[4] = s; r[5] = 0.f; r[6] = c; r[7] = 0.f;
    } else {  // diag(c - is, c + is)
      r[0] = c; r[1] = -s; r[2] = 0.f; r[3] = 0.f; r[4] = 0.f; r[5] = 0.f; r[6] = c; r[7] = s;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = r[j];
  }
}

// out[i] = in[i ^ tm] where every bit of cmask is set in i, else in[i]
__global__ void __launch_bounds__(kGateThreads)
k_mcx(float *out, const float *in, int64_t total, int64_t tm, int64_t cmask) {
  for (int64_t i = blockIdx.x * (int64_t)kGateThreads + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kGateThreads)
    st2(out, i, ld2(in, (i & cmask) == cmask ? (i ^ tm) : i));
}

// out[i] = -in[i] where every bit of mask is set in i (local index), else in[i]
__global__ void __launch_bounds__(kGateThreads)
k_phase_flip(float *out, const float *in, int64_t total, int64_t mask) {
  for (int64_t i = blockIdx.x * (int64_t)kGateThreads + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kGateThreads) {
    const float2 a = ld2(in, i);
    st2(out, i, (i & mask) == mask ? make_float2(-a.x, -a.y) : a);
  }
}

// out[i] = in[i with bits a and b exchanged] where every bit of cmask is set
__global__ void __launch_bounds__(kGateThreads)
k_swap(float *out, const float *in, int64_t total, int abit, int bbit, int64_t cmask) {
  const int64_t both = (int64_t(1) << abit) | (int64_t(1) << bbit);
  for (int64_t i = blockIdx.x * (int64_t)kGateThreads + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kGateThreads) {
    const bool differ = ((i >> abit) & 1) != ((i >> bbit) & 1);
    st2(out, i, ld2(in, (differ && (i & cmask) == cmask) ? (i ^ both) : i));
  }
}

// Reset of one bit, per pair (a0: bit 0, a1: bit 1):
//   a0' = a0 * N / (|a0| + 1e-7), a1' = 0,  N = sqrt(|a0|^2 + |a1|^2)
__global__ void __launch_bounds__(kGateThreads)
k_reset(float *out, const float *in, int64_t npairs, int bit) {
  const int64_t m = int64_t(1) << bit;
  for (int64_t p = blockIdx.x * (int64_t)kGateThreads + threadIdx.x; p < npairs;
       p += (int64_t)gridDim.x * kGateThreads) {
    const int64_t i0 = insert_zero(p, bit), i1 = i0 | m;
    const float2 a0 = ld2(in, i0), a1 = ld2(in, i1);
    const float r = sqrtf(a0.x * a0.x + a0.y * a0.y);
    const float N = sqrtf(a0.x * a0.x + a0.y * a0.y + a1.x * a1.x + a1.y * a1.y);
    const float s = N / (r + 1e-7f);
    st2(out, i0, make_float2(a0.x * s, a0.y * s));
    st2(out, i1, make_float2(0.f, 0.f));
  }
}

// Gradient of k_reset for a real loss (torch convention g = dL/dRe + i dL/dIm):
//   D = Re(conj(a0) g0);  gin0 = s g0 + D k0 a0,  gin1 = D a1 / (N (r + 1e-7))
//   k0 = ((r + 1e-7) / N - N / r) / (r + 1e-7)^2
// with the terms of a zero norm (r = 0 or N = 0) taken as 0.
__global__ void __launch_bounds__(kGateThreads)
k_reset_backward(float *gin, const float *g, const float *in, int64_t npairs, int bit) {
  const int64_t m = int64_t(1) << bit;
  for (int64_t p = blockIdx.x * (int64_t)kGateThreads + threadIdx.x; p < npairs;
       p += (int64_t)gridDim.x * kGateThreads) {
    const int64_t i0 = insert_zero(p, bit), i1 = i0 | m;
    const float2 a0 = ld2(in, i0), a1 = ld2(in, i1), g0 = ld2(g, i0);
    const float r = sqrtf(a0.x * a0.x + a0.y * a0.y);
    const float N = sqrtf(a0.x * a0.x + a0.y * a0.y + a1.x * a1.x + a1.y * a1.y);
    const float re = r + 1e-7f;
    const float s = N / re;
    const float D = a0.x * g0.x + a0.y * g0.y;
    const float k0 = (N > 0.f && r > 0.f) ? (re / N - N / r) / (re * re) : 0.f;
    const float k1 = N > 0.f ? D / (N * re) : 0.f;
    st2(gin, i0, make_float2(s * g0.x + D * k0 * a0.x, s * g0.y + D * k0 * a0.y));
    st2(gin, i1, make_float2(k1 * a1.x, k1 * a1.y));
  }
}

// ---------------------------------------------------------------- embedding / readout
// v = the first 2^n entries of (x[b, 0..d), pad, pad, ...); out[b] = v / max(||v||, 1e-12)
// when normalize, else v (imaginary parts 0).  One block per sample.
__device__ float gate_block_sum(float v, float *red) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < kGateThreads / 64; ++k) s += red[k];
  return s;
}

__device__ __forceinline__ float amp_entry(const float *xb, int d, int64_t i, float pad) {
  return i < d ? xb[i] : pad;
}

__global__ void __launch_bounds__(kGateThreads)
k_amp_embed(float *out, const float *x, int d, int n, int64_t x_stride, float pad, int normalize) {
  __shared__ float red[kGateThreads / 64];
  const int64_t b = blockIdx.x, N = int64_t(1) << n;
  const float *xb = x + b * x_stride;
  float scale = 1.f;
  if (normalize) {
    float ss = 0.f;
    for (int64_t i = threadIdx.x; i < N; i += kGateThreads) {
      const float v = amp_entry(xb, d, i, pad);
      ss += v * v;
    }
    scale = 1.f / fmaxf(sqrtf(gate_block_sum(ss, red)), 1e-12f);
  }
  for (int64_t i = threadIdx.x; i < N; i += kGateThreads)
    st2(out, b * N + i, make_float2(amp_entry(xb, d, i, pad) * scale, 0.f));
}

// gx[b, i] (i < min(d, 2^n)) from g = dL/dRe + i dL/dIm of the embedded state:
// without normalize gx = Re g; with it gx = Re g / r - v_i (v . Re g) / r^3, r = ||v||
// (r <= 1e-12: the clamped norm is constant, gx = Re g / 1e-12)
__global__ void __launch_bounds__(kGateThreads)
k_amp_embed_backward(float *gx, const float *g, const float *x, int d, int n, int64_t x_stride, float pad,
                     int normalize) {
  __shared__ float red[kGateThreads / 64];
  const int64_t b = blockIdx.x, N = int64_t(1) << n;
  const int dm = d < N ? d : (int)N;
  const float *xb = x + b * x_stride;
  float r = 1.f, k = 0.f;
  if (normalize) {
    float ss = 0.f, vg = 0.f;
    for (int64_t i = threadIdx.x; i < N; i += kGateThreads) {
      const float v = amp_entry(xb, d, i, pad);
      ss += v * v;
      vg += v * ld2(g, b * N + i).x;
    }
    const float sumsq = gate_block_sum(ss, red);
    const float dot = gate_block_sum(vg, red);
    const float nrm = sqrtf(sumsq);
    r = fmaxf(nrm, 1e-12f);
    k = nrm > 1e-12f ? dot / (nrm * nrm * nrm) : 0.f;
  }
  for (int64_t i = threadIdx.x; i < dm; i += kGateThreads) {
    const float gi = ld2(g, b * N + i).x;
    gx[b * d + i] = normalize ? gi / r - xb[i] * k : gi;
  }
  for (int64_t i = dm + threadIdx.x; i < d; i += kGateThreads) gx[b * d + i] = 0.f;  // cropped entries
}

// out[i] = |psi[i]|^2 (MeasureJointProbability) and its gradient gpsi = 2 g psi
__global__ void __launch_bounds__(kGateThreads) k_abs2(float *out, const float *psi, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)kGateThreads + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * kGateThreads) {
    const float2 a = ld2(psi, i);
    out[i] = a.x * a.x + a.y * a.y;
  }
}

__global__ void __launch_bounds__(kGateThreads)
k_abs2_backward(float *gpsi, const float *g, const float *psi, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)kGateThreads + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * kGateThreads) {
    const float2 a = ld2(psi, i);
    const float w = 2.f * g[i];
    st2(gpsi, i, make_float2(w * a.x, w * a.y));
  }
}

}  // namespace qs

using namespace qs;

static bool gate_shape_bad(int64_t B, int n) { return B < 1 || n < 1 || n > QS_MAX_QUBITS; }

static bool mask_ok(int64_t mask, int n) { return mask >= 0 && (n >= 63 || (mask >> n) == 0); }

extern "C" {

int qs_unitary_1q(float *out, const float *in, int64_t B, int n, int bit, const float *u, int64_t u_stride,
                  int64_t cmask, int dagger, void *stream) {
  if (gate_shape_bad(B, n) || !out || !in || !u || bit < 0 || bit >= n || u_stride < 0 || u_stride > 1 ||
      !mask_ok(cmask, n) || ((cmask >> bit) & 1))
    return QS_EINVAL;
  const int64_t npairs = (B << n) >> 1;
  hipLaunchKernelGGL(k_unitary_1q, dim3(gate_grid(npairs)), dim3(kGateThreads), 0, (hipStream_t)stream, out, in,
                     npairs, n, bit, u, u_stride, cmask, dagger);
  return launch_status();
}

int qs_rot_unitaries(float *u, const float *cs, int64_t count, int kind, void *stream) {
  if (count < 0 || kind < 0 || kind > 2 || (count > 0 && (!u || !cs))) return QS_EINVAL;
  if (count == 0) return QS_OK;
  hipLaunchKernelGGL(k_rot_unitaries, dim3(gate_grid(count)), dim3(kGateThreads), 0, (hipStream_t)stream, u, cs,
                     count, kind);
  return launch_status();
}

int qs_mcx(float *out, const float *in, int64_t B, int n, int tbit, int64_t cmask, void *stream) {
  if (gate_shape_bad(B, n) || !out || !in || out == in || tbit < 0 || tbit >= n || !mask_ok(cmask, n) ||
      ((cmask >> tbit) & 1))
    return QS_EINVAL;
  const int64_t total = B << n;
  hipLaunchKernelGGL(k_mcx, dim3(gate_grid(total)), dim3(kGateThreads), 0, (hipStream_t)stream, out, in, total,
                     int64_t(1) << tbit, cmask);
  return launch_status();
}

int qs_phase_flip(float *out, const float *in, int64_t B, int n, int64_t mask, void *stream) {
  if (gate_shape_bad(B, n) || !out || !in || mask == 0 || !mask_ok(mask, n)) return QS_EINVAL;
  const int64_t total = B << n;
  hipLaunchKernelGGL(k_phase_flip, dim3(gate_grid(total)), dim3(kGateThreads), 0, (hipStream_t)stream, out, in,
                     total, mask);
  return launch_status();
}

int qs_swap(float *out, const float *in, int64_t B, int n, int abit, int bbit, int64_t cmask, void *stream) {
  if (gate_shape_bad(B, n) || !out || !in || out == in || abit < 0 || abit >= n || bbit < 0 || bbit >= n ||
      !mask_ok(cmask, n) || ((cmask >> abit) & 1) || ((cmask >> bbit) & 1))
    return QS_EINVAL;
  const int64_t total = B << n;
  hipLaunchKernelGGL(k_swap, dim3(gate_grid(total)), dim3(kGateThreads), 0, (hipStream_t)stream, out, in, total,
                     abit, bbit, cmask);
  return launch_status();
}

int qs_reset(float *out, const float *in, int64_t B, int n, int bit, void *stream) {
  if (gate_shape_bad(B, n) || !out || !in || bit < 0 || bit >= n) return QS_EINVAL;
  const int64_t npairs = (B << n) >> 1;
  hipLaunchKernelGGL(k_reset, dim3(gate_grid(npairs)), dim3(kGateThreads), 0, (hipStream_t)stream, out, in, npairs,
                     bit);
  return launch_status();
}

int qs_reset_backward(float *gin, const float *g, const float *in, int64_t B, int n, int bit, void *stream) {
  if (gate_shape_bad(B, n) || !gin || !g || !in || gin == in || bit < 0 || bit >= n) return QS_EINVAL;
  const int64_t npairs = (B << n) >> 1;
  hipLaunchKernelGGL(k_reset_backward, dim3(gate_grid(npairs)), dim3(kGateThreads), 0, (hipStream_t)stream, gin, g,
                     in, npairs, bit);
  return launch_status();
}

int qs_amplitude_embed(float *out, const float *x, int64_t B, int d, int n, int64_t x_stride, float pad,
                       int normalize, void *stream) {
  if (gate_shape_bad(B, n) || !out || !x || d < 1 || x_stride < 0 || B > 0x7fffffff) return QS_EINVAL;
  hipLaunchKernelGGL(k_amp_embed, dim3((unsigned)B), dim3(kGateThreads), 0, (hipStream_t)stream, out, x, d, n,
                     x_stride, pad, normalize);
  return launch_status();
}

int qs_amplitude_embed_backward(float *gx, const float *g, const float *x, int64_t B, int d, int n,
                                int64_t x_stride, float pad, int normalize, void *stream) {
  if (gate_shape_bad(B, n) || !gx || !g || !x || d < 1 || x_stride < 0 || B > 0x7fffffff) return QS_EINVAL;
  hipLaunchKernelGGL(k_amp_embed_backward, dim3((unsigned)B), dim3(kGateThreads), 0, (hipStream_t)stream, gx, g,
                     x, d, n, x_stride, pad, normalize);
  return launch_status();
}

int qs_abs2(float *out, const float *psi, int64_t count, void *stream) {
  if (count < 0 || (count > 0 && (!out || !psi))) return QS_EINVAL;
  if (count == 0) return QS_OK;
  hipLaunchKernelGGL(k_abs2, dim3(gate_grid(count)), dim3(kGateThreads), 0, (hipStream_t)stream, out, psi, count);
  return launch_status();
}

int qs_abs2_backward(float *gpsi, const float *g, const float *psi, int64_t count, void *stream) {
  if (count < 0 || (count > 0 && (!gpsi || !g || !psi))) return QS_EINVAL;
  if (count == 0) return QS_OK;
  hipLaunchKernelGGL(k_abs2_backward, dim3(gate_grid(count)), dim3(kGateThreads), 0, (hipStream_t)stream, gpsi, g,
                     psi, count);
  return launch_status();
}

}  // extern "C"
