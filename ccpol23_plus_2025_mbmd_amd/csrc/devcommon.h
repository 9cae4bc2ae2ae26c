// Device helpers shared by the evaluation translation units.
#pragma once

#include "ccpol3b_internal.h"

namespace ccpol3b {

template <typename T> struct Pair2;
template <> struct Pair2<double> { using type = double2; };
template <> struct Pair2<float> { using type = float2; };

__device__ __forceinline__ double minimg_e(double d, double L) {
  return L > 0.0 ? d - L * round(d / L) : d;
}

// Minimum image and squared norm without FMA contraction: the same arithmetic
// as the list kernels, for acceptance tests that must agree with them bit for
// bit (d - L*nint(d/L); (x*x + y*y) + z*z).
__device__ __forceinline__ double minimg_nc(double d, double L) {
#pragma clang fp contract(off)
  return L > 0.0 ? d - L * round(d / L) : d;
}

__device__ __forceinline__ double sq3_nc(double x, double y, double z) {
#pragma clang fp contract(off)
  return x * x + y * y + z * z;
}

// A list outgrew its capacity in this call: the caller's outputs are not
// written and the call is repeated with larger capacities.
__device__ __forceinline__ bool lists_overflowed(const Dev& D) {
  return D.totals[T_NBOVER] > 0 || D.totals[T_N3] > D.cap3;
}

// owner of flat index s in an exclusive-offset array off[0..n]
__device__ __forceinline__ int owner_of(const int* off, int n, long s) {
  int lo = 0, hi = n;  // off[lo] <= s < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= s) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Appends the values of the flagged threads of a 256-thread block to q[base..]
// in thread order and returns how many there were.  Every thread of the block
// calls it; wcnt is a 4-int LDS scratch.  Ends with a barrier, so q is visible
// to the whole block on return.
__device__ __forceinline__ int block_append(bool flag, int val, int* q, int base, int* wcnt) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t m = __ballot(flag);
  if (lane == 0) wcnt[w] = __popcll(m);
  __syncthreads();
  int off = base;
  for (int u = 0; u < w; ++u) off += wcnt[u];
  if (flag) q[off + __popcll(m & lanemask_lt())] = val;
  const int tot = (wcnt[0] + wcnt[1]) + (wcnt[2] + wcnt[3]);
  __syncthreads();
  return tot;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Deterministic block sum over 256 threads (4 wavefronts).
__device__ __forceinline__ double block_sum256(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// Segmented inclusive sum over runs of equal keys within a wavefront (keys
// form contiguous runs); the last lane of each run ends up with the run total.
__device__ __forceinline__ double seg_sum(double v, int key, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double vu = __shfl_up(v, d, 64);
    const int ku = __shfl_up(key, d, 64);
    if (lane >= d && ku == key) v += vu;
  }
  return v;
}

template <typename T> __device__ __forceinline__ T t_exp(T x);
template <> __device__ __forceinline__ double t_exp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float t_exp<float>(float x) { return __expf(x); }

// x**delta for the damping exponent: exponentiation by squaring when delta is
// a small whole number (10 in both shipped CONTROL files), pow otherwise.
__device__ __forceinline__ double pow_delta(double x, const Consts& K) {
  if (K.idelta > 0) {
    double r = 1.0, b = x;
    for (int e = K.idelta; e; e >>= 1) {
      if (e & 1) r *= b;
      b *= b;
    }
    return r;
  }
  return pow(x, K.delta33);
}

}  // namespace ccpol3b
