// Shared device helpers for the gfx950 state-vector kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qandle_sv.h"

namespace qs {

// last HIP error seen by a launch wrapper (reported through qs_last_hip_error)
extern int g_last_hip_error;

inline int launch_status() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    g_last_hip_error = (int)e;
    return QS_EHIP;
  }
  return QS_OK;
}

// Grid for a kernel that walks its work items in a grid-stride loop: every
// workgroup that can be resident on the current device at once (from the
// occupancy query, cached per kernel, block size and device), capped by the
// number of items.  Host-side; allocates nothing.
template <typename Kernel>
unsigned resident_grid(Kernel kernel, int block, int64_t items) {
  static int cached[64];  // per device: resident workgroups + 1 (0 = not queried)
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  int g = cached[dev] - 1;
  if (g <= 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu < 1)
      per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
    g = per_cu * cus;
    cached[dev] = g + 1;
  }
  if (items < 1) return 1;
  return (unsigned)(items < g ? items : g);
}

__device__ __forceinline__ float2 ld2(const float *p, int64_t i) {
  return reinterpret_cast<const float2 *>(p)[i];
}
__device__ __forceinline__ void st2(float *p, int64_t i, float2 v) {
  reinterpret_cast<float2 *>(p)[i] = v;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Rotation R_P(phi) = exp(-i phi P / 2) on one amplitude pair (a0: bit = 0,
// a1: bit = 1), c = cos(phi/2), s = sin(phi/2).  Column convention
// psi' = U psi, the transpose of the reference's row-vector `state @ M`
// (src/qandle/operators.py:236-238, 292).  The inverse uses -s.
template <int KIND>
__device__ __forceinline__ void rot_pair(float c, float s, float2 &a0, float2 &a1) {
  float2 n0, n1;
  if (KIND == QS_RX) {  // [[c, -is], [-is, c]]
    n0 = make_float2(c * a0.x + s * a1.y, c * a0.y - s * a1.x);
    n1 = make_float2(c * a1.x + s * a0.y, c * a1.y - s * a0.x);
  } else if (KIND == QS_RY) {  // [[c, -s], [s, c]]
    n0 = make_float2(c * a0.x - s * a1.x, c * a0.y - s * a1.y);
    n1 = make_float2(s * a0.x + c * a1.x, s * a0.y + c * a1.y);
  } else {  // diag(c - is, c + is)
    n0 = make_float2(c * a0.x + s * a0.y, c * a0.y - s * a0.x);
    n1 = make_float2(c * a1.x - s * a1.y, c * a1.y + s * a1.x);
  }
  a0 = n0;
  a1 = n1;
}

__device__ __forceinline__ void rot_pair_rt(int kind, float c, float s, float2 &a0, float2 &a1) {
  if (kind == QS_RX)
    rot_pair<QS_RX>(c, s, a0, a1);
  else if (kind == QS_RY)
    rot_pair<QS_RY>(c, s, a0, a1);
  else
    rot_pair<QS_RZ>(c, s, a0, a1);
}

// Im( conj(l0) (P psi)_0 + conj(l1) (P psi)_1 ) for one pair.
template <int KIND>
__device__ __forceinline__ float pauli_im_pair(float2 l0, float2 l1, float2 p0, float2 p1) {
  // conj(a) b = (a.x b.x + a.y b.y) + i (a.x b.y - a.y b.x)
  if (KIND == QS_RX) {
    return (l0.x * p1.y - l0.y * p1.x) + (l1.x * p0.y - l1.y * p0.x);
  } else if (KIND == QS_RY) {  // (P psi)_0 = -i psi1, (P psi)_1 = i psi0
    return (l1.x * p0.x + l1.y * p0.y) - (l0.x * p1.x + l0.y * p1.y);
  } else {
    return (l0.x * p0.y - l0.y * p0.x) - (l1.x * p1.y - l1.y * p1.x);
  }
}

// index with a zero bit inserted at position `bit`
__device__ __forceinline__ int64_t insert_zero(int64_t p, int bit) {
  int64_t low = p & ((int64_t(1) << bit) - 1);
  return ((p >> bit) << (bit + 1)) | low;
}

}  // namespace qs
