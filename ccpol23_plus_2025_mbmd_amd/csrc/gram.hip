// Gram matrix G = Phi^T Phi of the coefficient-gradient design matrix
// (rows = configurations, columns = the 728 SR/LR linear coefficients) on the
// FP32 matrix cores, with FP64 accuracy.
//
// Every column a is scaled by its largest magnitude, 2^e_a, and written as six
// base-2^9 digits in FP32: Phi[r,a] = 2^e_a sum_k d_k[r,a] 2^(-9(k+1)), |d_k| <= 2^9.
// Digit products are integers below 2^18, so a 64-row block of products sums
// to an integer below 2^24: v_mfma_f32_16x16x4f32 accumulates it exactly.
// Each 64-row block is then scaled and added in FP64.  Digit pairs with
// k + l <= 5 are kept (21 pairs); the dropped ones are below 2^-54 of
// 2^(e_a+e_b) per row, the FP64 level.
#include "devcommon.h"

namespace ccpol3b {

constexpr int GR_DIG = 6;       // digits per value
constexpr int GR_BITS = 9;      // bits per digit
constexpr int GR_ROWBLK = 64;   // rows accumulated exactly in FP32
constexpr int GR_TILE = 32;     // output tile per 256-thread block (2 x 2 wavefronts of 16 x 16)

// Column exponents: e_a with max_r |Phi[r,a]| < 2^e_a (0 for an all-zero column).
__global__ __launch_bounds__(256) void k_gram_exps(int nrow, int ncol, const double* __restrict__ phi, int ld,
                                                   int* __restrict__ ex) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ncol) return;
  double m = 0.0;
  for (int r = 0; r < nrow; ++r) m = fmax(m, fabs(phi[(long)r * ld + a]));
  int e = 0;
  if (m > 0.0) frexp(m, &e);  // m = f 2^e, f in [0.5, 1)
  ex[a] = e;
}

// Digits dig[k][r][a] (rows padded to a multiple of 64, columns to 32, padding zero).
__global__ __launch_bounds__(256) void k_gram_digits(int nrow, int ncol, int nrp, int ncp,
                                                     const double* __restrict__ phi, int ld,
                                                     const int* __restrict__ ex, float* __restrict__ dig) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nrp * ncp) return;
  const int r = (int)(i / ncp), a = (int)(i - (long)r * ncp);
  double x = (r < nrow && a < ncol) ? ldexp(phi[(long)r * ld + a], -ex[a]) : 0.0;
  const long plane = (long)nrp * ncp;
#pragma unroll
  for (int k = 0; k < GR_DIG; ++k) {
    const double y = ldexp(x, GR_BITS);
    const double d = rint(y);
    x = y - d;  // exact
    dig[k * plane + i] = (float)d;
  }
}

// One block per upper-triangle tile (ta <= tb) of 32 x 32 entries; wavefront w
// owns the 16 x 16 sub-tile (w / 2, w % 2).  Lane layout of
// v_mfma_f32_16x16x4f32: A[m][kk] from lane m + 16 kk, B[kk][n] from lane
// n + 16 kk, D[4 (lane / 16) + i][lane % 16] in result register i.
__global__ __launch_bounds__(256) void k_gram_mfma(int ncol, int nrp, int ncp, const float* __restrict__ dig,
                                                   const int* __restrict__ ex, double* __restrict__ gram,
                                                   int ldg) {
  // tile (ta, tb) from the linear upper-triangle index
  const int nt = ncp / GR_TILE;
  int t = blockIdx.x, ta = 0;
  while (t >= nt - ta) {
    t -= nt - ta;
    ++ta;
  }
  const int tb = ta + t;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int a0 = ta * GR_TILE + (w >> 1) * 16, b0 = tb * GR_TILE + (w & 1) * 16;
  const int col = lane & 15, kk = lane >> 4;
  const long plane = (long)nrp * ncp;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < GR_DIG; ++k) {
    for (int l = 0; k + l < GR_DIG; ++l) {
      const double sc = ldexp(1.0, -GR_BITS * (k + l + 2));
      const float* A = dig + k * plane + a0 + col;
      const float* B = dig + l * plane + b0 + col;
      for (int r0 = 0; r0 < nrp; r0 += GR_ROWBLK) {
        typedef float v4f __attribute__((ext_vector_type(4)));
        v4f c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < GR_ROWBLK / 4; ++s) {
          const long row = r0 + 4 * s + kk;
          c = __builtin_amdgcn_mfma_f32_16x16x4f32(A[row * ncp], B[row * ncp], c, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] += sc * (double)c[i];
      }
    }
  }
  const int b = b0 + col;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int a = a0 + 4 * kk + i;
    if (a < ncol && b < ncol && a <= b) {
      const double v = ldexp(acc[i], ex[a] + ex[b]);
      gram[(long)a * ldg + b] = v;
      gram[(long)b * ldg + a] = v;
    }
  }
}

}  // namespace ccpol3b
