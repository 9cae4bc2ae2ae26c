// Gram matrix G = Phi^T Phi of the coefficient-gradient design matrix
// (rows = configurations, columns = the 728 SR/LR linear coefficients) on the
// FP64 matrix cores (v_mfma_f64_16x16x4f64): products and sums in FP64, so
// the result carries the ordinary FP64 dot-product error bound.
//
// Work split: 64 x 64 output tiles of the upper triangle (the lower one is the
// mirror) times `nsplit` row ranges, one 256-thread block each, so a 4096 x 728
// matrix gives 78 x 8 blocks.  Each block stages 16 rows of its two column
// strips in LDS per step; wavefront w owns the 32 x 32 quarter (w / 2, w % 2)
// as 2 x 2 MFMA tiles.  Row-range partial tiles go to a scratch array and a
// second kernel adds them in a fixed order (results are the same run to run).
#include "devcommon.h"

namespace ccpol3b {

constexpr int GR_T = 64;   // output tile edge
constexpr int GR_K = 16;   // rows per LDS stage
constexpr int GR_LD = GR_T + 1;

typedef double v4d __attribute__((ext_vector_type(4)));

// upper-triangle tile (ta <= tb) of the linear index t over nt x nt tiles
__device__ __forceinline__ void gram_tile(int t, int nt, int& ta, int& tb) {
  ta = 0;
  while (t >= nt - ta) {
    t -= nt - ta;
    ++ta;
  }
  tb = ta + t;
}

// part[s][a][b] = sum over rows r of split s of phi[r][a] phi[r][b], for (a, b)
// in the tile.  Lane layout of v_mfma_f64_16x16x4f64 (measured on gfx950):
// A[m][k] from lane m + 16 k, B[k][n] from lane n + 16 k, and result element
// i of a lane is D[lane / 16 + 4 i][lane % 16] (rows interleaved, unlike the
// four consecutive rows of v_mfma_f32_16x16x4f32).
__global__ __launch_bounds__(256) void k_gram_f64(int nrow, int ncol, const double* __restrict__ phi, int ld,
                                                  int rows_per_split, double* __restrict__ part) {
  __shared__ double sA[GR_K][GR_LD];
  __shared__ double sB[GR_K][GR_LD];
  const int nt = (ncol + GR_T - 1) / GR_T;
  int ta, tb;
  gram_tile(blockIdx.x, nt, ta, tb);
  const int s = blockIdx.y;
  const int r0 = s * rows_per_split, r1 = min(nrow, r0 + rows_per_split);
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int a0 = ta * GR_T, b0 = tb * GR_T;
  const int m = lane & 15, k4 = lane >> 4;
  const int wa = (w >> 1) * 32, wb = (w & 1) * 32;
  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
  // staging: thread t loads column (t % 64) of rows (t / 64) + 4 q, q = 0..3
  const int lc = tid & 63, lr = tid >> 6;
  for (int rb = r0; rb < r1; rb += GR_K) {
#pragma unroll
    for (int q = 0; q < GR_K / 4; ++q) {
      const int r = rb + lr + 4 * q;
      const bool rin = r < r1;
      const int ca = a0 + lc, cb = b0 + lc;
      sA[lr + 4 * q][lc] = (rin && ca < ncol) ? phi[(long)r * ld + ca] : 0.0;
      sB[lr + 4 * q][lc] = (rin && cb < ncol) ? phi[(long)r * ld + cb] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GR_K / 4; ++kk) {
      const int row = 4 * kk + k4;
      const double av0 = sA[row][wa + m], av1 = sA[row][wa + 16 + m];
      const double bv0 = sB[row][wb + m], bv1 = sB[row][wb + 16 + m];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av0, bv0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av0, bv1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av1, bv0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av1, bv1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
  double* out = part + (long)s * ncol * ncol;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int a = a0 + wa + 16 * i + k4 + 4 * e, b = b0 + wb + 16 * j + m;
        if (a < ncol && b < ncol) out[(long)a * ncol + b] = acc[i][j][e];
      }
}

// gram[a][b] = gram[b][a] = sum over splits, in order, of part[s][a][b] (a <= b).
__global__ __launch_bounds__(256) void k_gram_sum(int ncol, int nsplit, const double* __restrict__ part,
                                                  double* __restrict__ gram, int ldg) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)ncol * ncol) return;
  const int a = (int)(i / ncol), b = (int)(i - (long)a * ncol);
  if (a > b) return;
  double v = 0.0;
  for (int s = 0; s < nsplit; ++s) v += part[(long)s * ncol * ncol + i];
  gram[(long)a * ldg + b] = v;
  gram[(long)b * ldg + a] = v;
}

}  // namespace ccpol3b
