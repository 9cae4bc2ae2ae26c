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

// A list outgrew its capacity in this call (the call is repeated with larger
// capacities), or a force fell outside the fixed-point range (the call fails):
// either way the caller's outputs are not written.
__device__ __forceinline__ bool lists_overflowed(const Dev& D) {
  return D.totals[T_NBOVER] > 0 || D.totals[T_N3] > D.cap3 || D.totals[T_N3] > D.grid3 || D.totals[T_FXOVER] != 0;
}

// Forces are summed in 128-bit fixed point, in units of 2^-96 hartree/bohr,
// each accumulator a pair of 64-bit words (lo, hi) with value
// hi * 2^-32 + lo * 2^-96.  A contribution is added to lo with one integer
// atomic, and hi receives its own part plus the carry out of lo.  Integer
// addition is associative and the number of carries out of lo depends only on
// the sum, not on the order of the additions, so every sum is exact and the
// same bits come out whatever order blocks add in: every run, and a box
// evaluated alone or inside a batch.  Contributions are rounded down to
// 2^-96 hartree/bohr (1.3e-29) and must stay below 2^30 in magnitude; a NaN
// or an infinite contribution (overlapping sites) flags the call instead.
constexpr double FX_LIMIT = 1073741824.0;  // 2^30

struct Fx128 {
  unsigned long long lo;
  unsigned long long hi;  // two's complement
};

__device__ __forceinline__ Fx128 fx_split(double v) {
  const double a = v * 4294967296.0;  // in units of 2^-32
  const double h = floor(a);
  Fx128 r;
  r.hi = (unsigned long long)(long long)h;
  r.lo = (unsigned long long)((a - h) * 18446744073709551616.0);  // the fraction in units of 2^-96
  return r;
}

// w[0..1] += x (global or LDS words), exactly.
__device__ __forceinline__ void fx_atomic_add(unsigned long long* w, Fx128 x) {
  const unsigned long long old = atomicAdd(w, x.lo);
  const unsigned long long carry = old + x.lo < old ? 1ull : 0ull;
  atomicAdd(w + 1, x.hi + carry);
}

__device__ __forceinline__ double fx_decode(const unsigned long long* w) {
  return (double)(long long)w[1] * (1.0 / 4294967296.0) + (double)w[0] * 0x1p-96;
}

// v in fixed point; a NaN or out-of-range v flags the call and counts as zero.
__device__ __forceinline__ Fx128 fx_checked(const Dev& D, double v) {
  if (!(fabs(v) < FX_LIMIT)) {
    atomicOr(D.totals + T_FXOVER, 1);
    return Fx128{0ull, 0ull};
  }
  return fx_split(v);
}

// accumulator idx += v (hartree/bohr).
__device__ __forceinline__ void fx_add(const Dev& D, long idx, double v) {
  fx_atomic_add(D.acc + 2 * idx, fx_checked(D, v));
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

// 1/sqrt(x) and 1/x in FP64 for positive normal x: the hardware estimates
// (v_rsq_f64, v_rcp_f64) refined by two Newton-Raphson steps each,
// y <- y + y (1 - x y^2) / 2 and y <- y + y (1 - x y); a few ulp, against
// ~27 and ~12 instructions for the IEEE 1/sqrt(x) and division.
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double r = fma(-(x * y), y, 1.0);
  y = fma(0.5 * y, r, y);
  r = fma(-(x * y), y, 1.0);
  return fma(0.5 * y, r, y);
}

__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over the 64 lanes of a wavefront, valid in lane 63 only: DPP row shifts
// build each row's sum in its lane 15, then row broadcasts add the rows.  A
// fixed order of additions (the same bits on every run), with no LDS traffic.
__device__ __forceinline__ double dpp_step(double v, int ctrl_id) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  int a, b;
  switch (ctrl_id) {  // row_shr:1, 2, 4, 8 within rows of 16; row_bcast15 into rows 1, 3; row_bcast31 into rows 2, 3
    case 0: a = __builtin_amdgcn_update_dpp(0, lo, 0x111, 0xf, 0xf, true); b = __builtin_amdgcn_update_dpp(0, hi, 0x111, 0xf, 0xf, true); break;
    case 1: a = __builtin_amdgcn_update_dpp(0, lo, 0x112, 0xf, 0xf, true); b = __builtin_amdgcn_update_dpp(0, hi, 0x112, 0xf, 0xf, true); break;
    case 2: a = __builtin_amdgcn_update_dpp(0, lo, 0x114, 0xf, 0xf, true); b = __builtin_amdgcn_update_dpp(0, hi, 0x114, 0xf, 0xf, true); break;
    case 3: a = __builtin_amdgcn_update_dpp(0, lo, 0x118, 0xf, 0xf, true); b = __builtin_amdgcn_update_dpp(0, hi, 0x118, 0xf, 0xf, true); break;
    case 4: a = __builtin_amdgcn_update_dpp(0, lo, 0x142, 0xa, 0xf, false); b = __builtin_amdgcn_update_dpp(0, hi, 0x142, 0xa, 0xf, false); break;
    default: a = __builtin_amdgcn_update_dpp(0, lo, 0x143, 0xc, 0xf, false); b = __builtin_amdgcn_update_dpp(0, hi, 0x143, 0xc, 0xf, false); break;
  }
  return v + __hiloint2double(b, a);
}
__device__ __forceinline__ double wave_sum_lane63(double v) {
#pragma unroll
  for (int k = 0; k < 6; ++k) v = dpp_step(v, k);
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

// Geometry of an FS3 triplet from its COM vectors x_ij, x_ik and the derived
// x_jk (:862-884): distances, R_avg (:894-898), the damping f_d and its
// derivative (:900-904), and the SR/LR weights of the cubic switch (:1420-1423).
__device__ __forceinline__ void tri_geometry(const double* cij, const double* cik, const double* cjk,
                                             const Consts& K, int mi, int mj, int mk, TriGeo& g) {
  for (int c = 0; c < 3; ++c) {
    g.c[0][c] = cij[c];
    g.c[1][c] = cik[c];
    g.c[2][c] = cjk[c];
  }
  for (int b = 0; b < 3; ++b) g.R[b] = sqrt(g.c[b][0] * g.c[b][0] + g.c[b][1] * g.c[b][1] + g.c[b][2] * g.c[b][2]);
  const double Ravg = (g.R[0] + g.R[2] + g.R[1]) / 3.0;
  const double pom = exp(K.alp3 * (Ravg - K.r0dmp3));
  const double pp = 1.0 / (1.0 + pom);
  const double fd = pow_delta(pp, K);
  g.Ravg = Ravg;
  g.fd = fd;
  g.dfd = -K.alp3 * K.delta33 * pom * pp * fd;  // d f_d / d R_avg
  g.w[0] = 1.0; g.wp[0] = 0.0; g.w[1] = 0.0; g.wp[1] = 0.0;
  if (Ravg <= K.rmin) {
    g.nreg = 1; g.reg0 = 0;
  } else if (Ravg >= K.rmax) {
    g.nreg = 1; g.reg0 = 1;
  } else {
    const double span = K.rmax - K.rmin;
    const double c = (Ravg - K.rmin) / span;
    const double sm = 1.0 + 2.0 * c * c * c - 3.0 * c * c;
    const double dsm = (6.0 * c * c - 6.0 * c) / span;
    g.nreg = 2; g.reg0 = 0;
    g.w[0] = sm; g.wp[0] = dsm;
    g.w[1] = 1.0 - sm; g.wp[1] = -dsm;
  }
  g.m[0] = mi; g.m[1] = mj; g.m[2] = mk;
  g.pad = 0;
}

// e^x in FP64 for the FS3 pair functions (arguments of magnitude below ~700
// there; beyond that the result still saturates to 0 or inf through ldexp).
// Cody-Waite reduction x = n ln2 + r, |r| <= ln2/2, with ln2 split into a head
// of 32 significant bits (n head is exact for |n| < 2^21) and a tail; e^r by
// the degree-11 polynomial that interpolates it at the 12 Chebyshev nodes of
// [-ln2/2, ln2/2] (interpolation error 4e-18 relative; coefficients computed in
// 80-bit arithmetic and rounded to double), in Horner form; the result is
// 2^n e^r.  Measured against an extended-precision exp: at most 1.6 ulp.
__device__ __forceinline__ double exp_cw(double x) {
  constexpr double LN2_HI = 0x1.62e42fee00000p-1;
  constexpr double LN2_LO = 0x1.a39ef35800000p-33;
  constexpr double C[12] = {
    0x1.0000000000000p+0,
    0x1.0000000000000p+0,
    0x1.0000000000010p-1,
    0x1.5555555555558p-3,
    0x1.555555554f11fp-5,
    0x1.111111110f69dp-7,
    0x1.6c16c187efa7ap-10,
    0x1.a01a01b04adf3p-13,
    0x1.a01991c5c4583p-16,
    0x1.71ddf6fafee26p-19,
    0x1.28b3f2d37e699p-22,
    0x1.af615189357d7p-26};
  const double n = __builtin_rint(x * 1.4426950408889634);
  const double r = fma(-n, LN2_LO, fma(-n, LN2_HI, x));
  double p = C[11];
#pragma unroll
  for (int k = 10; k >= 0; --k) p = fma(p, r, C[k]);
  return __builtin_ldexp(p, (int)n);
}

template <typename T> __device__ __forceinline__ T t_exp(T x);
template <> __device__ __forceinline__ double t_exp<double>(double x) { return exp_cw(x); }
template <> __device__ __forceinline__ float t_exp<float>(float x) { return __expf(x); }


}  // namespace ccpol3b
