// Geometry preparation, neighbour lists and the compacted FS3/FS2 lists.
//
// Every acceptance predicate reproduces the reference's FP64 arithmetic bit
// for bit: positions are converted to bohr by division (threebody_potential.f90
// :48-62), minimum images are d - L*nint(d/L) with NINT rounding half away from
// zero, squares are summed left to right, and this file is compiled without
// FMA contraction (the reference's x86-64 build has no FMA).
//
// Work decomposition: one block of four wavefronts per molecule for the
// neighbour search (which also counts the molecule's FS3 triplets) and the FS3
// list fill, one wavefront per
// directed neighbour pair for the exported FS2 items.  The neighbour search
// tests the molecules of the 27 surrounding cells of a cell grid when the box
// has one (periodic, at least CELL_MIN cells a side of at least the larger
// cutoff), and every molecule of the box otherwise; both produce the same
// lists, in ascending molecule index.
// Lists come out in the reference's loop orders: FS3 (i,j,k) lexicographic
// with i<j<k (:859-860), FS2 (k,i,j) with i<j (:297-301).
#pragma clang fp contract(off)

#include "devcommon.h"

namespace ccpol3b {

__device__ __forceinline__ double minimg(double d, double L) {
  return L > 0.0 ? d - L * round(d / L) : d;
}

__device__ __forceinline__ double sq3(double x, double y, double z) {
  return x * x + y * y + z * z;  // (x*x + y*y) + z*z, no contraction
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

// ------------------------------------------------------------------ cell grids
// Cell of coordinate x (bohr) along one side: the position wrapped into [0, L)
// over the cell side L/nc.  Rounding can only move a molecule lying on a cell
// face into the adjacent cell, and the cell side exceeds the cutoff by a
// relative 1e-9, so two molecules within the cutoff (minimum image) are
// always in the same or adjacent cells.
__device__ __forceinline__ int cell_coord(double x, double L, int nc) {
  const double u = x - L * floor(x / L);
  const int q = (int)(u / (L / nc));
  return q < 0 ? 0 : (q >= nc ? nc - 1 : q);
}

// ------------------------------------------------------------------ prep
// COM (A) -> bohr and site offsets (A, Fortran order xs(is,i)) -> [m][s][c] bohr;
// zero the force accumulators and the counters of the call; each molecule's
// cell (or -1 when its box has no grid) and the cell occupancy counts, which
// k_cells_build leaves at zero for the next call.
__global__ void k_prep(int nm, int ld, const double* __restrict__ x, const double* __restrict__ y,
                       const double* __restrict__ z, const double* __restrict__ xs,
                       const double* __restrict__ ys, const double* __restrict__ zs, Dev D) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nm) {
    const double cx = x[t] / BOHR2A, cy = y[t] / BOHR2A, cz = z[t] / BOHR2A;
    D.cx[t] = cx;
    D.cy[t] = cy;
    D.cz[t] = cz;
    if (D.ncells > 0) {
      const int b = D.mol_box[t], nc = D.cell_nc[b];
      int cc = -1;
      if (nc >= CELL_MIN) {
        const double L = D.boxL[b];
        cc = D.cell0[b] + (cell_coord(cx, L, nc) * nc + cell_coord(cy, L, nc)) * nc + cell_coord(cz, L, nc);
        atomicAdd(D.cell_cnt + cc, 1);
      }
      D.mol_cell[t] = cc;
    }
  }
  if (t < nm * NS3) {
    const int m = t / NS3, s = t - m * NS3;
    const long src = s + (long)ld * m;
    double* dst = D.site + (long)t * 3;
    dst[0] = xs[src] / BOHR2A;
    dst[1] = ys[src] / BOHR2A;
    dst[2] = zs[src] / BOHR2A;
  }
  if (t < nm * ACC_W) D.acc[2 * t] = D.acc[2 * t + 1] = 0ull;
  if (t < T_COUNT) D.totals[t] = 0;
}

// ------------------------------------------------------------------ FS3 triplet sweep
// The FS3 triplets of molecule i are the pairs (j, k), j < k, of its upper FS3
// neighbours (j > i, ascending from the neighbour sweep), in the reference's
// order (j, then k: :859-860), accepted when the derived vector
// x_jk = dx_jk + L*n_ij - L*n_ik (:882-884) is inside the cutoff.  The upper
// neighbours are staged in LDS, 6 doubles per slot: position and integer images
// n_ij = nint((x_i - x_j)/L).
size_t fs3_list_lds_bytes(int cap_nb) { return (size_t)cap_nb * 6 * sizeof(double); }

__device__ __forceinline__ void stage_upper(double* st, int cap, int a, double xi, double yi, double zi, double xj,
                                            double yj, double zj, double L) {
  st[a] = xj;
  st[cap + a] = yj;
  st[2 * cap + a] = zj;
  st[3 * cap + a] = L > 0.0 ? round((xi - xj) / L) : 0.0;
  st[4 * cap + a] = L > 0.0 ? round((yi - yj) / L) : 0.0;
  st[5 * cap + a] = L > 0.0 ? round((zi - zj) / L) : 0.0;
}

// pair (r, s), r < s < m, of flat index t in row-major order over r
__device__ __forceinline__ void pair_of(long t, int m, int& r, int& s) {
  const double b = 2.0 * m - 1.0;
  int q = (int)((b - sqrt(b * b - 8.0 * (double)t)) * 0.5);  // rows before t: S_r = r m - r (r+1) / 2
  q = q < 0 ? 0 : (q > m - 2 ? m - 2 : q);
  while (q > 0 && (long)q * m - (long)q * (q + 1) / 2 > t) --q;
  while (q < m - 2 && (long)(q + 1) * m - (long)(q + 1) * (q + 2) / 2 <= t) ++q;
  r = q;
  s = q + 1 + (int)(t - ((long)q * m - (long)q * (q + 1) / 2));
}

// Appends in block order: the LB_WAVES wavefronts of a block each pass K
// wave-uniform counts (ballot popcounts); `before` gets the sum over lower
// wavefronts and `total` the block's sum, the same in every thread.  Entries
// then go to running offset + before + rank in the wavefront, so a round of
// LB_WAVES * 64 candidates keeps candidate order.  Two barriers; buf is LDS.
constexpr int LB_WAVES = 4;
constexpr int LB_THREADS = 64 * LB_WAVES;
template <int K>
__device__ __forceinline__ void wave_prefix(const int (&cnt)[K], int (&before)[K], int (&total)[K],
                                            int (*buf)[K]) {
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) buf[w][k] = cnt[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    before[k] = 0;
    total[k] = 0;
  }
#pragma unroll
  for (int u = 0; u < LB_WAVES; ++u)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      before[k] += u < w ? buf[u][k] : 0;
      total[k] += buf[u][k];
    }
  __syncthreads();
}

// The block sweeps the m staged neighbours of i, LB_THREADS pairs at a time.
// mode 0 returns the number of triplets; mode 1 also writes them from `out`
// with their geometry (TriGeo) and counts their regions in shist.
template <int mode>
__device__ int fs3_sweep(const Dev& D, const Consts& K, int i, const int* nb, int m, const double* st, int cap,
                         double L, double xi, double yi, double zi, long out, int* shist) {
  __shared__ int pbuf[LB_WAVES][1];
  const int lane = threadIdx.x & 63;
  const double *px = st, *py = st + cap, *pz = st + 2 * cap;
  const double *qx = st + 3 * cap, *qy = st + 4 * cap, *qz = st + 5 * cap;
  const long npair = (long)m * (m - 1) / 2;
  int count = 0;
  for (long t0 = 0; t0 < npair; t0 += LB_THREADS) {
    const long t = t0 + threadIdx.x;
    bool ok = false;
    int r = 0, s = 0;
    double xjk = 0.0, yjk = 0.0, zjk = 0.0;
    if (t < npair) {
      pair_of(t, m, r, s);
      xjk = px[r] - px[s];
      yjk = py[r] - py[s];
      zjk = pz[r] - pz[s];
      if (L > 0.0) {
        xjk = xjk + L * qx[r] - L * qx[s];
        yjk = yjk + L * qy[r] - L * qy[s];
        zjk = zjk + L * qz[r] - L * qz[s];
      }
      ok = sq3(xjk, yjk, zjk) < K.rc2_3;
    }
    const uint64_t msk = __ballot(ok);
    const int c[1] = {(int)__popcll(msk)};
    int before[1], total[1];
    wave_prefix<1>(c, before, total, pbuf);
    if (mode == 1 && ok) {
      const long p = out + count + before[0] + __popcll(msk & lanemask_lt());
      if (p < D.cap3) {
        D.tri[p] = make_int4(i, nb[r], nb[s], 0);
        const double cij[3] = {(xi - px[r]) - L * qx[r], (yi - py[r]) - L * qy[r], (zi - pz[r]) - L * qz[r]};
        const double cik[3] = {(xi - px[s]) - L * qx[s], (yi - py[s]) - L * qy[s], (zi - pz[s]) - L * qz[s]};
        const double cjk[3] = {xjk, yjk, zjk};
        TriGeo g;
        tri_geometry(cij, cik, cjk, K, i, nb[r], nb[s], g);
        D.geo3[p] = g;
        atomicAdd(shist + (g.nreg == 2 ? 1 : 2 * g.reg0), 1);
      }
    }
    (void)lane;
    count += total[0];
  }
  return count;
}

// ------------------------------------------------------------------ neighbour lists
// One block (four wavefronts) per molecule i; the candidates j are the
// molecules of the 27 cells around i's cell when the box has a cell grid
// (:297-336 test every molecule; the cells leave out only molecules beyond
// the cutoffs), every other molecule of the box otherwise.  Each is tested
// with the reference's pair predicate (min image of x_i - x_j), both cutoffs
// at once.  The sweep meets j in ascending order; the cell path ranks the
// accepted j and writes each at its rank, so both give the same lists.  With
// do3 the upper FS3 neighbours are staged in LDS on the way and the block then
// counts i's FS3 triplets into cnt3s[i], so that the triplet offsets come out of
// the scan that follows and the list can be filled in one more launch.
__global__ __launch_bounds__(LB_THREADS) void k_neighbours(Dev D, Consts K, int do3) {
  extern __shared__ double lsm[];
  __shared__ int pbuf[LB_WAVES][4];
  const int tid = threadIdx.x, lane = tid & 63;
  const int i = blockIdx.x;
  if (i >= D.nm) return;  // the whole block
  const int b = D.mol_box[i];
  const double L = D.boxL[b];
  const int j0 = D.box_start[b], j1 = D.box_start[b + 1];
  const double xi = D.cx[i], yi = D.cy[i], zi = D.cz[i];
  const int cap = D.cap_nb;
  int n3 = 0, n2 = 0, u3 = 0, u2 = 0;  // running counts, the same in every thread
  int* nb3 = D.nb3 + (long)i * cap;
  int* nb2 = D.nb2 + (long)i * cap;
  double* v2 = D.nbv2 + (long)i * 3 * cap;
  const int nc = D.cell_nc[b];
  if (nc >= CELL_MIN) {
    // candidates: the molecules of the 27 cells around i's cell (distinct, as
    // nc >= 3), cell q's members at flat positions [spre[q], spre[q+1])
    __shared__ int spre[28], sbeg[27];
    int* A3 = reinterpret_cast<int*>(lsm + (do3 ? 6 * cap : 0));  // accepted, in candidate order
    int* A2 = A3 + cap;
    if (tid < 64) {
      const int c0 = D.cell0[b], ci = D.mol_cell[i] - c0;
      const int ix = ci / (nc * nc), iy = (ci / nc) % nc, iz = ci % nc;
      int len = 0;
      if (lane < 27) {
        int jx = ix + lane / 9 - 1, jy = iy + (lane / 3) % 3 - 1, jz = iz + lane % 3 - 1;
        jx += jx < 0 ? nc : (jx >= nc ? -nc : 0);
        jy += jy < 0 ? nc : (jy >= nc ? -nc : 0);
        jz += jz < 0 ? nc : (jz >= nc ? -nc : 0);
        const int cc = c0 + (jx * nc + jy) * nc + jz;
        sbeg[lane] = D.cell_start[cc];
        len = D.cell_start[cc + 1] - sbeg[lane];
      }
      int pre = len;  // inclusive prefix over lanes 0..26
#pragma unroll
      for (int d = 1; d < 32; d <<= 1) {
        const int u = __shfl_up(pre, d, 64);
        if (lane >= d) pre += u;
      }
      if (lane < 27) spre[lane + 1] = pre;
      if (lane == 0) spre[0] = 0;
    }
    __syncthreads();
    const int total = spre[27];
    for (int f0 = 0; f0 < total; f0 += LB_THREADS) {
      const int f = f0 + tid;
      bool in3 = false, in2 = false;
      int j = -1;
      if (f < total) {
        int lo = 0, hi = 27;  // spre[lo] <= f < spre[hi]
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (spre[mid] <= f) lo = mid; else hi = mid;
        }
        j = D.cell_mol[sbeg[lo] + f - spre[lo]];
        if (j != i) {
          const double d2 = sq3(minimg(xi - D.cx[j], L), minimg(yi - D.cy[j], L), minimg(zi - D.cz[j], L));
          in3 = d2 < K.rc2_3;
          in2 = d2 < K.rc2_2;
        }
      }
      const uint64_t m3 = __ballot(in3), m2 = __ballot(in2);
      const int c[4] = {(int)__popcll(m3), (int)__popcll(m2), (int)__popcll(__ballot(in3 && j < i)),
                        (int)__popcll(__ballot(in2 && j < i))};
      int before[4], tot[4];
      wave_prefix<4>(c, before, tot, pbuf);
      if (in3) {
        const int p = n3 + before[0] + __popcll(m3 & lanemask_lt());
        if (p < cap) A3[p] = j;
      }
      if (in2) {
        const int p = n2 + before[1] + __popcll(m2 & lanemask_lt());
        if (p < cap) A2[p] = j;
      }
      n3 += tot[0];
      n2 += tot[1];
      u3 += tot[2];
      u2 += tot[3];
    }
    __syncthreads();
    // ascending order: each stored entry goes to its rank among the stored entries
    const int s3 = n3 < cap ? n3 : cap, s2 = n2 < cap ? n2 : cap;
    for (int e = tid; e < s3; e += LB_THREADS) {
      const int j = A3[e];
      int r = 0;
      for (int f = 0; f < s3; ++f) r += A3[f] < j ? 1 : 0;
      nb3[r] = j;
      const int pu = r - u3;  // position among the upper entries
      if (do3 && j > i && pu >= 0 && pu < cap) stage_upper(lsm, cap, pu, xi, yi, zi, D.cx[j], D.cy[j], D.cz[j], L);
    }
    for (int e = tid; e < s2; e += LB_THREADS) {
      const int j = A2[e];
      int r = 0;
      for (int f = 0; f < s2; ++f) r += A2[f] < j ? 1 : 0;
      nb2[r] = j;
      v2[r] = minimg(xi - D.cx[j], L);
      v2[cap + r] = minimg(yi - D.cy[j], L);
      v2[2 * cap + r] = minimg(zi - D.cz[j], L);
    }
  } else {
    for (int base = j0; base < j1; base += LB_THREADS) {
      const int j = base + tid;
      bool in3 = false, in2 = false;
      double xj = 0.0, yj = 0.0, zj = 0.0, dx = 0.0, dy = 0.0, dz = 0.0;
      if (j < j1 && j != i) {
        xj = D.cx[j];
        yj = D.cy[j];
        zj = D.cz[j];
        dx = minimg(xi - xj, L);
        dy = minimg(yi - yj, L);
        dz = minimg(zi - zj, L);
        const double d2 = sq3(dx, dy, dz);
        in3 = d2 < K.rc2_3;
        in2 = d2 < K.rc2_2;
      }
      const uint64_t m3 = __ballot(in3), m2 = __ballot(in2);
      const int c[4] = {(int)__popcll(m3), (int)__popcll(m2), (int)__popcll(__ballot(in3 && j < i)),
                        (int)__popcll(__ballot(in2 && j < i))};
      int before[4], tot[4];
      wave_prefix<4>(c, before, tot, pbuf);
      if (in3) {
        const int p = n3 + before[0] + __popcll(m3 & lanemask_lt());
        if (p < cap) nb3[p] = j;
        // upper entries (j > i) are staged at their position among the upper
        // ones; the candidates ascend, so every lower entry of the round precedes them
        const int pu = p - u3 - tot[2];
        if (do3 && j > i && pu < cap) stage_upper(lsm, cap, pu, xi, yi, zi, xj, yj, zj, L);
      }
      if (in2) {
        const int p = n2 + before[1] + __popcll(m2 & lanemask_lt());
        if (p < cap) {
          nb2[p] = j;
          v2[p] = dx;
          v2[cap + p] = dy;
          v2[2 * cap + p] = dz;
        }
      }
      n3 += tot[0];
      n2 += tot[1];
      u3 += tot[2];
      u2 += tot[3];
    }
  }
  const int nn3 = n3 < cap ? n3 : cap, nn2 = n2 < cap ? n2 : cap;
  const int up3 = u3 < nn3 ? u3 : nn3, up2 = u2 < nn2 ? u2 : nn2;
  if (tid == 0) {
    if (n3 > cap || n2 > cap) atomicMax(D.totals + T_NBOVER, n3 > n2 ? n3 : n2);
    D.nb3_cnt[i] = nn3;
    D.nb2_cnt[i] = nn2;
    D.up3[i] = up3;  // first upper position (entries < i come first)
    D.up2[i] = up2;
    D.sl2_cnt[i] = nn2;        // FS2 sub-lists (k; i) over all neighbours i
    D.pcnt2[i] = nn2 - up2;    // FS2 pairs i<j
  }
  if (do3) {
    __syncthreads();  // the staged neighbours
    const int count = fs3_sweep<0>(D, K, i, nullptr, nn3 - up3, lsm, cap, L, xi, yi, zi, 0, nullptr);
    if (tid == 0) D.cnt3s[i] = count;
  }
}


// ------------------------------------------------------------------ FS3 triplets
// One block (four wavefronts) per molecule i: stages its upper FS3
// neighbours again and writes its triplets from off3s[i] (fs3_sweep), with the
// geometry of each triplet and the SR / switch / LR histogram.
__global__ __launch_bounds__(LB_THREADS) void k_fs3_fill(Dev D, Consts K) {
  extern __shared__ double lsm[];
  __shared__ int shist[3];
  const int lane = threadIdx.x;
  const int i = blockIdx.x;
  if (i >= D.nm) return;  // the whole block
  if (lane < 3) shist[lane] = 0;
  const int cap = D.cap_nb;
  const int n = D.nb3_cnt[i], up = D.up3[i], m = n - up;
  const int* nb = D.nb3 + (long)i * cap + up;
  const double L = D.boxL[D.mol_box[i]];
  const double xi = D.cx[i], yi = D.cy[i], zi = D.cz[i];
  for (int a = lane; a < m; a += LB_THREADS) {
    const int j = nb[a];
    stage_upper(lsm, cap, a, xi, yi, zi, D.cx[j], D.cy[j], D.cz[j], L);
  }
  __syncthreads();
  fs3_sweep<1>(D, K, i, nb, m, lsm, cap, L, xi, yi, zi, D.off3s[i], shist);
  // one global atomic per molecule and region
  __syncthreads();
  if (lane < 3 && shist[lane] > 0) atomicAdd(D.totals + T_HIST + lane, shist[lane]);
}

// ------------------------------------------------------------------ FS2 items
// Ordered items (k; i<j): one wavefront per sub-list (k; i = nb2(k)[a]); the
// candidates j are the upper entries of nb2(i) other than k, accepted when the
// derived k-j vector x_kj = x_ki + x_ij is inside the cutoff (:325-336).
__global__ __launch_bounds__(256) void k_fs2_items(Dev D, Consts K, int mode) {
  const int lane = threadIdx.x & 63;
  const long nsl = D.totals[T_SL2];
  const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
  for (long s = (long)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); s < nsl; s += nwaves) {
  const int k = owner_of(D.sl2_off, D.nm, s);
  const int a = (int)(s - D.sl2_off[k]);
  const int i = D.nb2[(long)k * D.cap_nb + a];
  const double L = D.boxL[D.mol_box[k]];
  const double xi = D.cx[i], yi = D.cy[i], zi = D.cz[i];
  const double xki = minimg(D.cx[k] - xi, L), yki = minimg(D.cy[k] - yi, L), zki = minimg(D.cz[k] - zi, L);
  const int* nbi = D.nb2 + (long)i * D.cap_nb;
  const int ni = D.nb2_cnt[i];
  const int u0 = D.up2[i];
  const long pbase = D.poff2[i];
  const long out = mode ? D.off2s[s] : 0;
  int count = 0;
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
  if (mode == 0 && lane == 0) D.cnt2s[s] = count;
  }
}

// ------------------------------------------------------------------ scans
// Exclusive scan of n ints by one block of NT threads, tile by tile through LDS
// (coalesced loads, 8 consecutive elements per thread); off[n] and *total get
// the sum.
// Exclusive scans of K int arrays of length n together, by one block of NT
// threads: off[k][q] = sum of cnt[k][0..q-1], and off[k][n] and *total[k]
// receive the sums; the last array's sum is also returned to every thread.
// off may alias cnt (each tile is read before it is written).  Tiles are
// staged in buf (SCAN_BUF ints shared by the K arrays).
constexpr int SCAN_BUF = 8192;

template <int NT, int K>
__device__ long block_exclusive_scan(int* buf, const int* const (&cnt)[K], int* const (&off)[K], long n,
                                     int* const (&total)[K]) {
  constexpr int PER = SCAN_BUF / (K * NT);  // entries per thread and array
  constexpr int TILE = NT * PER;
  static_assert(PER >= 1, "scan buffer too small");
  __shared__ long wsum[K][NT / 64];
  __shared__ long carry_s[K];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t < K) carry_s[t] = 0;
  __syncthreads();
  for (long base = 0; base < n; base += TILE) {
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int r = 0; r < PER; ++r) {
        const long q = base + r * NT + t;
        buf[k * TILE + r * NT + t] = q < n ? cnt[k][q] : 0;
      }
    __syncthreads();
    long loc[K][PER], s[K], incl[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      s[k] = 0;
#pragma unroll
      for (int r = 0; r < PER; ++r) {
        loc[k][r] = s[k];
        s[k] += buf[k * TILE + t * PER + r];
      }
      incl[k] = s[k];
    }
    // inclusive scans of the per-thread sums: wavefront shuffles, then wavefront totals
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const long u = __shfl_up(incl[k], d, 64);
        if (lane >= d) incl[k] += u;
      }
    if (lane == 63)
#pragma unroll
      for (int k = 0; k < K; ++k) wsum[k][w] = incl[k];
    __syncthreads();
    long excl[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      long wbase = 0;
      for (int u = 0; u < w; ++u) wbase += wsum[k][u];
      excl[k] = carry_s[k] + wbase + incl[k] - s[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
      for (int r = 0; r < PER; ++r) buf[k * TILE + t * PER + r] = (int)(excl[k] + loc[k][r]);
      if (t == NT - 1) carry_s[k] = excl[k] + s[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int r = 0; r < PER; ++r) {
        const long q = base + r * NT + t;
        if (q < n) off[k][q] = buf[k * TILE + r * NT + t];
      }
    __syncthreads();
  }
  if (t == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      off[k][n] = (int)carry_s[k];
      *total[k] = (int)carry_s[k];
    }
  const long last = carry_s[K - 1];
  __syncthreads();
  return last;
}

// One block: exclusive scan of the cell counts into cell_start, then every
// molecule into its cell's range of cell_mol (order within a cell is not
// used: the neighbour search ranks what it accepts).  Counting each cell down
// leaves cell_cnt at zero, as k_prep of the next call expects.
template <int NT>
__global__ __launch_bounds__(NT) void k_cells_build(Dev D) {
  __shared__ int buf[SCAN_BUF];
  const int* const c[1] = {D.cell_cnt};
  int* const o[1] = {D.cell_start};
  int* const tt[1] = {D.cell_start + D.ncells};
  block_exclusive_scan<NT, 1>(buf, c, o, D.ncells, tt);
  for (int m = threadIdx.x; m < D.nm; m += NT) {
    const int cc = D.mol_cell[m];
    if (cc >= 0) D.cell_mol[D.cell_start[cc] + atomicSub(D.cell_cnt + cc, 1) - 1] = m;
  }
}
template __global__ void k_cells_build<1024>(Dev);

// Triplet range of each box (the list is sorted by its first molecule, and
// boxes are contiguous molecule ranges) and the number of reduction chunks of
// each box, for FS3 triplets and for the FS2 energies (fs2_parts entries per
// molecule); the chunk counts go to box_c3/box_c2 and are scanned in place.
// Called by a whole block after the FS3 offsets are final.
__device__ void box_ranges(const Dev& D, int do3, int do2) {
  for (int b = threadIdx.x; b <= D.nbox; b += blockDim.x) {
    const int m = D.box_start[b];
    const int t3 = do3 ? (m < D.nm ? D.off3s[m] : D.totals[T_N3]) : 0;
    const int t2 = do2 ? m * D.fs2_parts : 0;
    D.box_t3[b] = t3;
    D.box_t2[b] = t2;
    int c3 = 0, c2 = 0;
    if (b < D.nbox) {
      const int m1 = D.box_start[b + 1];
      const int e3 = do3 ? (m1 < D.nm ? D.off3s[m1] : D.totals[T_N3]) : 0;
      const int e2 = do2 ? m1 * D.fs2_parts : 0;
      c3 = (e3 - t3 + RED_CHUNK - 1) / RED_CHUNK;
      c2 = (e2 - t2 + RED_CHUNK - 1) / RED_CHUNK;
    }
    D.box_c3[b] = c3;
    D.box_c2[b] = c2;
  }
  __syncthreads();
}

// phase 0: per-molecule counts -> FS2 sub-list and pair offsets, FS3 triplet
//          offsets; then the per-box triplet ranges and the first reduction
//          chunk of each box
// phase 2: per-sub-list candidate counts -> FS2 item offsets (list export only)
template <int NT>
__global__ __launch_bounds__(NT) void k_scan(Dev D, int phase, int do3, int do2) {
  __shared__ int buf[SCAN_BUF];
  int* const T = D.totals;
  if (phase == 0) {
    // the per-molecule scans of both terms together
    long n3 = 0;
    if (do2 && do3) {
      const int* const c[3] = {D.sl2_cnt, D.pcnt2, D.cnt3s};
      int* const o[3] = {D.sl2_off, D.poff2, D.off3s};
      int* const tt[3] = {T + T_SL2, T + T_NP2, T + T_N3};
      n3 = block_exclusive_scan<NT, 3>(buf, c, o, D.nm, tt);
    } else if (do2) {
      const int* const c[2] = {D.sl2_cnt, D.pcnt2};
      int* const o[2] = {D.sl2_off, D.poff2};
      int* const tt[2] = {T + T_SL2, T + T_NP2};
      block_exclusive_scan<NT, 2>(buf, c, o, D.nm, tt);
    } else if (do3) {
      const int* const c[1] = {D.cnt3s};
      int* const o[1] = {D.off3s};
      int* const tt[1] = {T + T_N3};
      n3 = block_exclusive_scan<NT, 1>(buf, c, o, D.nm, tt);
    }
    if (D.nbox == 1) {
      // one box (molecules [0, nm)): its ranges and reduction chunks follow
      // from the totals, without reading the offsets back
      if (threadIdx.x == 0) {
        const int n2 = do2 ? D.nm * D.fs2_parts : 0;
        const int c3 = (int)((n3 + RED_CHUNK - 1) / RED_CHUNK), c2 = (n2 + RED_CHUNK - 1) / RED_CHUNK;
        D.box_t3[0] = 0;
        D.box_t3[1] = (int)n3;
        D.box_t2[0] = 0;
        D.box_t2[1] = n2;
        D.box_c3[0] = 0;
        D.box_c3[1] = c3;
        D.box_c2[0] = 0;
        D.box_c2[1] = c2;
        T[T_NCH3] = c3;
        T[T_NCH2] = c2;
      }
      return;
    }
    box_ranges(D, do3, do2);
    const int* const c[2] = {D.box_c3, D.box_c2};
    int* const o[2] = {D.box_c3, D.box_c2};
    int* const tt[2] = {T + T_NCH3, T + T_NCH2};
    block_exclusive_scan<NT, 2>(buf, c, o, D.nbox, tt);
  } else if (phase == 2) {
    if (do2) {
      const int* const c[1] = {D.cnt2s};
      int* const o[1] = {D.off2s};
      int* const tt[1] = {T + T_NI2L};
      block_exclusive_scan<NT, 1>(buf, c, o, D.totals[T_SL2], tt);
    }
  }
}
template __global__ void k_scan<1024>(Dev, int, int, int);
template __global__ void k_scan<256>(Dev, int, int, int);

}  // namespace ccpol3b
