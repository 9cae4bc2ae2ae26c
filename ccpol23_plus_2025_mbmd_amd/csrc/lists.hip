// Geometry preparation, neighbour lists and the compacted FS3/FS2 lists.
//
// Every acceptance predicate reproduces the reference's FP64 arithmetic bit
// for bit: positions are converted to bohr by division (threebody_potential.f90
// :48-62), minimum images are d - L*nint(d/L) with NINT rounding half away from
// zero, squares are summed left to right, and this file is compiled without
// FMA contraction (the reference's x86-64 build has no FMA).
#pragma clang fp contract(off)

#include "ccpol3b_internal.h"

namespace ccpol3b {

__device__ __forceinline__ double minimg(double d, double L) {
  return L > 0.0 ? d - L * round(d / L) : d;
}

__device__ __forceinline__ double sq3(double x, double y, double z) {
  return x * x + y * y + z * z;  // (x*x + y*y) + z*z, no contraction
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// ------------------------------------------------------------------ prep
// COM (A) -> bohr and site offsets (A, Fortran order xs(is,i)) -> [m][s][c] bohr.
__global__ void k_prep(int nm, int ld, const double* __restrict__ x, const double* __restrict__ y,
                       const double* __restrict__ z, const double* __restrict__ xs,
                       const double* __restrict__ ys, const double* __restrict__ zs, Dev D) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nm) {
    D.cx[t] = x[t] / BOHR2A;
    D.cy[t] = y[t] / BOHR2A;
    D.cz[t] = z[t] / BOHR2A;
  }
  if (t < nm * NS3) {
    const int m = t / NS3, s = t - m * NS3;
    const long src = s + (long)ld * m;
    double* dst = D.site + (long)t * 3;
    dst[0] = xs[src] / BOHR2A;
    dst[1] = ys[src] / BOHR2A;
    dst[2] = zs[src] / BOHR2A;
  }
}

// ------------------------------------------------------------------ neighbour lists
// One wavefront per molecule i; every other molecule j of the same box is
// tested with the reference's pair predicate (min image of x_i - x_j).  Both
// cutoffs are produced in one sweep; lists are in ascending j.
__global__ __launch_bounds__(256) void k_neighbours(Dev D, Consts K) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= D.nm) return;
  const int i = wave;
  const int b = D.mol_box[i];
  const double L = D.boxL[b];
  const int j0 = D.box_start[b], j1 = D.box_start[b + 1];
  const double xi = D.cx[i], yi = D.cy[i], zi = D.cz[i];
  int n3 = 0, n2 = 0;
  int* nb3 = D.nb3 + (long)i * D.cap_nb;
  int* nb2 = D.nb2 + (long)i * D.cap_nb;
  for (int base = j0; base < j1; base += 64) {
    const int j = base + lane;
    bool in3 = false, in2 = false;
    if (j < j1 && j != i) {
      const double dx = minimg(xi - D.cx[j], L);
      const double dy = minimg(yi - D.cy[j], L);
      const double dz = minimg(zi - D.cz[j], L);
      const double d2 = sq3(dx, dy, dz);
      in3 = d2 < K.rc2_3;
      in2 = d2 < K.rc2_2;
    }
    const uint64_t m3 = __ballot(in3), m2 = __ballot(in2);
    if (in3) {
      const int p = n3 + __popcll(m3 & lanemask_lt());
      if (p < D.cap_nb) nb3[p] = j;
    }
    if (in2) {
      const int p = n2 + __popcll(m2 & lanemask_lt());
      if (p < D.cap_nb) nb2[p] = j;
    }
    n3 += __popcll(m3);
    n2 += __popcll(m2);
  }
  if (lane == 0) {
    if (n3 > D.cap_nb || n2 > D.cap_nb) atomicMax(D.totals + 3, n3 > n2 ? n3 : n2);
    D.nb3_cnt[i] = n3 < D.cap_nb ? n3 : D.cap_nb;
    D.nb2_cnt[i] = n2 < D.cap_nb ? n2 : D.cap_nb;
  }
}

// first position in an ascending list whose entry exceeds `v`
__device__ __forceinline__ int upper_start(const int* list, int n, int v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (list[mid] > v) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// ------------------------------------------------------------------ FS3 triplets
// One wavefront per i: candidates (j,k) with i<j<k both in nb3(i) (so the i-j and
// i-k minimum-image tests already passed); the j-k test uses the reference's
// derived vector x_jk = dx_jk + L*nint(dx_ij/L) - L*nint(dx_ik/L) (:882-884).
// mode 0 counts into cnt3[i]; mode 1 writes at off3[i] in (j,k) lexicographic order.
__global__ __launch_bounds__(256) void k_fs3_list(Dev D, Consts K, int mode) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= D.nm) return;
  const int i = wave;
  const int* nb = D.nb3 + (long)i * D.cap_nb;
  const int n = D.nb3_cnt[i];
  const int u0 = upper_start(nb, n, i);
  const double L = D.boxL[D.mol_box[i]];
  const double xi = D.cx[i], yi = D.cy[i], zi = D.cz[i];
  long out = mode ? D.off3[i] : 0;
  int count = 0;
  for (int a = u0; a < n - 1; ++a) {
    const int j = nb[a];
    const double xj = D.cx[j], yj = D.cy[j], zj = D.cz[j];
    double nxij = 0.0, nyij = 0.0, nzij = 0.0;
    if (L > 0.0) {
      nxij = round((xi - xj) / L);
      nyij = round((yi - yj) / L);
      nzij = round((zi - zj) / L);
    }
    for (int bb = a + 1; bb < n; bb += 64) {
      const int q = bb + lane;
      bool ok = false;
      int k = 0;
      if (q < n) {
        k = nb[q];
        const double xk = D.cx[k], yk = D.cy[k], zk = D.cz[k];
        double xjk = xj - xk, yjk = yj - yk, zjk = zj - zk;
        if (L > 0.0) {
          const double nxik = round((xi - xk) / L);
          const double nyik = round((yi - yk) / L);
          const double nzik = round((zi - zk) / L);
          xjk = xjk + L * nxij - L * nxik;
          yjk = yjk + L * nyij - L * nyik;
          zjk = zjk + L * nzij - L * nzik;
        }
        ok = sq3(xjk, yjk, zjk) < K.rc2_3;
      }
      const uint64_t m = __ballot(ok);
      if (mode && ok) {
        const long p = out + count + __popcll(m & lanemask_lt());
        if (p < D.cap3) D.tri[p] = make_int4(i, j, k, 0);
      }
      count += __popcll(m);
    }
  }
  if (mode == 0 && lane == 0) D.cnt3[i] = count;
}

// ------------------------------------------------------------------ FS2 pairs and items
// Unordered pairs i<j within rcut3: the upper part of nb2(i).
__global__ void k_fs2_pairs(Dev D, int mode) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= D.nm) return;
  const int* nb = D.nb2 + (long)i * D.cap_nb;
  const int n = D.nb2_cnt[i];
  const int u0 = upper_start(nb, n, i);
  if (mode == 0) {
    D.pcnt2[i] = n - u0;
  } else {
    const long o = D.poff2[i];
    for (int q = u0; q < n; ++q)
      if (o + q - u0 < D.cap_pairs2) D.pairs2[o + q - u0] = make_int2(i, nb[q]);
  }
}

// Ordered items (k; i<j) in the reference loop order k -> i -> j (:297-301):
// i in nb2(k), j in the upper part of nb2(i), j != k, and the derived k-j
// vector x_kj = x_ki + x_ij inside the cutoff (:325-336).  One wavefront per k.
__global__ __launch_bounds__(256) void k_fs2_items(Dev D, Consts K, int mode) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= D.nm) return;
  const int k = wave;
  const int* nbk = D.nb2 + (long)k * D.cap_nb;
  const int nk = D.nb2_cnt[k];
  const double L = D.boxL[D.mol_box[k]];
  const double xk = D.cx[k], yk = D.cy[k], zk = D.cz[k];
  long out = mode ? D.off2[k] : 0;
  int count = 0;
  for (int a = 0; a < nk; ++a) {
    const int i = nbk[a];
    const double xi = D.cx[i], yi = D.cy[i], zi = D.cz[i];
    const double xki = minimg(xk - xi, L), yki = minimg(yk - yi, L), zki = minimg(zk - zi, L);
    const int* nbi = D.nb2 + (long)i * D.cap_nb;
    const int ni = D.nb2_cnt[i];
    const int u0 = upper_start(nbi, ni, i);
    const long pbase = D.poff2[i];
    for (int bb = u0; bb < ni; bb += 64) {
      const int q = bb + lane;
      bool ok = false;
      int j = 0;
      if (q < ni) {
        j = nbi[q];
        if (j != k) {
          const double xij = minimg(xi - D.cx[j], L);
          const double yij = minimg(yi - D.cy[j], L);
          const double zij = minimg(zi - D.cz[j], L);
          ok = sq3(xki + xij, yki + yij, zki + zij) < K.rc2_2;
        }
      }
      const uint64_t m = __ballot(ok);
      if (mode && ok) {
        const long p = out + count + __popcll(m & lanemask_lt());
        if (p < D.cap_items2) D.items2[p] = make_int4(k, i, j, (int)(pbase + q - u0));
      }
      count += __popcll(m);
    }
  }
  if (mode == 0 && lane == 0) D.cnt2[k] = count;
}

// ------------------------------------------------------------------ scans
// Exclusive scans of the three per-molecule counts (one block of 1024 threads).
__device__ void block_exclusive_scan(const int* cnt, int* off, int n, int* total) {
  __shared__ long part[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b0 = t * per, b1 = min(n, b0 + per);
  long s = 0;
  for (int q = b0; q < b1; ++q) s += cnt[q];
  part[t] = s;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    long v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  long run = part[t] - s;
  for (int q = b0; q < b1; ++q) {
    off[q] = (int)run;
    run += cnt[q];
  }
  if (t == 1023) {
    off[n] = (int)part[1023];
    *total = (int)part[1023];
  }
  __syncthreads();
}

__global__ __launch_bounds__(1024) void k_scan(Dev D, int which) {
  if (which & 1) block_exclusive_scan(D.cnt3, D.off3, D.nm, D.totals + 0);
  if (which & 2) block_exclusive_scan(D.pcnt2, D.poff2, D.nm, D.totals + 1);
  if (which & 4) block_exclusive_scan(D.cnt2, D.off2, D.nm, D.totals + 2);
}

}  // namespace ccpol3b
