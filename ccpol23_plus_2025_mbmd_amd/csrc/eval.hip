// FS3 evaluation kernels, the accumulator write-back and the per-box
// energy/virial reductions.
//
// FS3 (threebody_potential.f90:676-2062).  For an accepted triplet (i,j,k) and
// region X (SR or LR) the reference energy is
//   E_X = sum_{a in i, b in j, c in k} e(ab) e(ac) e(bc) P_X(r_bc, r_ac, r_ab)
// with e(r) = exp(-beta r)/(1+exp(-gamma (r-r0))) and P trilinear, whose
// coefficients depend on the site types through the orbit map.  Writing
// v_ab = (e(ab), e(ab) r_ab), E_X is trilinear in the three pair vectors:
//   E_X = sum C[(ta,p),(tb,q),(tc,s)] v_bc[p] v_ac[q] v_ab[s],
// C being the fully symmetric dense table of the 364 coefficients.  The
// gradient with respect to the pair vector of (x,y) summed over the third
// site z of type tz is
//   G[x,y,s] = sum_{tz,p,q} C[(tx,p),(ty,q),(tz,s)] M[x,y,tz,p,q],
//   M[x,y,tz,p,q] = sum_{z in tz} v_yz[p] v_xz[q],
// evaluated once per molecule pair role (three passes), so each site pair owns
// its gradient and no cross-thread accumulation is needed.  dE/dr follows from
// dv0/dr = e' and dv1/dr = e + r e', e' = e (-beta + gamma (1 - damp)).
// The triplet energy is f_d(R_avg) * sum_X w_X(R_avg) E_X with the damping and
// switch of :894-904 and :1420-1423.
#include <type_traits>

#include "devcommon.h"

namespace ccpol3b {

// First site of each type block (O | H H | D1 D1 | N1 x4 | N2 x4 | N3 x4).
constexpr int kTypeStart[NTYPE + 1] = {0, 1, 3, 5, 9, 13, 17};

// Operand layout of pass P (0: X=i,Y=j,Z=k; 1: X=i,Y=k,Z=j; 2: X=j,Y=k,Z=i) in
// the site-pair blocks (0 ij, 1 ik, 2 jk, each [first][second] row major):
// v_xz is A = block ablk, row of x at x*arow, next z at +az; v_yz is B = block
// bblk, row of y at y*brow, next z at +bz.
//   P=0: v_xz = ik[x][z], v_yz = jk[y][z]
//   P=1: v_xz = ij[x][z], v_yz = jk[z][y]
//   P=2: v_xz = ij[z][x], v_yz = ik[z][y]
struct PassLayout {
  int ablk, arow, az, bblk, brow, bz;
};
__host__ __device__ constexpr PassLayout pass_layout(int P) {
  return P == 0 ? PassLayout{1, NS3, 1, 2, NS3, 1}
       : P == 1 ? PassLayout{0, NS3, 1, 2, 1, NS3}
                : PassLayout{0, 1, NS3, 1, 1, NS3};
}

// One tile of NU x NW site pairs of a pass: NU sites u of one molecule (rows
// U + a URS), NW sites w of the other (rows W + b WRS), contracted over the 17
// sites z of the third molecule (steps UZS, WZS) and the 6 type blocks tz:
//   M[a][b][p][q] = sum_{z in tz} v_wz[p] v_uz[q],
//   G[a][b][s]   += sum_pq C[tu][tw][tz][p][q][s] M[a][b][p][q].
// Both loops are unrolled, so with compile-time strides every LDS read is a
// base register plus an immediate offset.  With dc (PHI), also
//   dE/dC[tu][tw][tz][p][q][s] += sum_ab vown[a][b][s] M[a][b][p][q].
// cp points at the tile's coefficients: C + (tu NTYPE + tw) NTYPE 8 in the
// dense layout, C + 2 (tu NTYPE + tw) in the interleaved one (CIL, cil_index).
template <typename T, int NU, int NW, bool PHI, bool CIL = false>
__device__ __forceinline__ void contract_tile(const typename Pair2<T>::type* U, int URS, int UZS,
                                              const typename Pair2<T>::type* W, int WRS, int WZS,
                                              const T* __restrict__ cp, T (&G)[NU][NW][2], T* dc,
                                              const T (&vown)[NU][NW][2]) {
  using T2 = typename Pair2<T>::type;
#pragma unroll
  for (int a = 0; a < NU; ++a)
#pragma unroll
    for (int b = 0; b < NW; ++b) G[a][b][0] = G[a][b][1] = T(0);
#pragma unroll 1
  for (int tz = 0; tz < NTYPE; ++tz) {
    // M over the type block's sites; every block holds at least one site, so
    // the first one initialises M
    T M[NU][NW][2][2];
    T uv[NU][2], wv[NW][2];
    // the type block's coefficients: the FP32 kernels issue the loads before
    // the site loop, which hides their latency behind it (the FP64 kernels have
    // no registers to spare for that and load them after it)
    T cc[8];
    auto load_c = [&]() {
#pragma unroll
      for (int k = 0; k < 8; ++k) cc[k] = CIL ? cp[cil_index(tz, k)] : cp[tz * 8 + k];
    };
    if constexpr (sizeof(T) == 4) load_c();
    auto load = [&]() {
#pragma unroll
      for (int a = 0; a < NU; ++a) {
        const T2 v = U[a * URS];
        uv[a][0] = v.x;
        uv[a][1] = v.y;
      }
#pragma unroll
      for (int b = 0; b < NW; ++b) {
        const T2 v = W[b * WRS];
        wv[b][0] = v.x;
        wv[b][1] = v.y;
      }
      U += UZS;
      W += WZS;
    };
    load();
#pragma unroll
    for (int a = 0; a < NU; ++a)
#pragma unroll
      for (int b = 0; b < NW; ++b)
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int q = 0; q < 2; ++q) M[a][b][p][q] = wv[b][p] * uv[a][q];
    const int nz = kTypeStart[tz + 1] - kTypeStart[tz];
#pragma unroll 1
    for (int z = 1; z < nz; ++z) {
      load();
#pragma unroll
      for (int a = 0; a < NU; ++a)
#pragma unroll
        for (int b = 0; b < NW; ++b)
#pragma unroll
          for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) M[a][b][p][q] += wv[b][p] * uv[a][q];
    }
    if constexpr (PHI) {
      if (dc) {
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int s = 0; s < 2; ++s) {
              T d = T(0);
#pragma unroll
              for (int a = 0; a < NU; ++a)
#pragma unroll
                for (int b = 0; b < NW; ++b) d += vown[a][b][s] * M[a][b][p][q];
              atomicAdd(dc + tz * 8 + (p * 2 + q) * 2 + s, d);
            }
      }
    }
    if constexpr (sizeof(T) == 8) load_c();
#pragma unroll
    for (int a = 0; a < NU; ++a)
#pragma unroll
      for (int b = 0; b < NW; ++b)
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int s = 0; s < 2; ++s) G[a][b][s] += cc[(p * 2 + q) * 2 + s] * M[a][b][p][q];
  }
}

// The triple-stage work of one thread.  Waves 0-2 own passes 0-2 on the site
// pairs of sites 1-16 (an 8 x 8 grid of 2 x 2 tiles aligned to the type
// blocks, strides known at compile time).  Wave 3 owns the O row and the O
// column of all three passes: lane 17 P + r, r < 8 the O row pairs (0, 2r+1)
// and (0, 2r+2); 8 <= r < 16 the O column pairs (2r-15, 0) and (2r-14, 0),
// evaluated with the roles of the two molecules swapped (U the O site of Y, W
// two sites of X: C is symmetric under that swap); r = 16 the pair (0, 0).
// own[a][b] is the site-pair index (in the pass's own block, ij / ik / jk) of
// tile pair (a, b), -1 when unused.
struct TileSpec {
  int P, own[2][2];
  int u_off, u_rs, u_zs, w_off, w_rs, w_zs, tu, tw;
  bool used;
};

__device__ __forceinline__ TileSpec tile_spec(int tid) {
  TileSpec t{};
  const int wave = tid >> 6, lane = tid & 63;
  if (wave < 3) {
    const PassLayout L = pass_layout(wave);
    const int x0 = 1 + 2 * (lane >> 3), y0 = 1 + 2 * (lane & 7);
    t.P = wave;
    t.used = true;
    t.u_off = L.ablk * NPAIR + x0 * L.arow; t.u_rs = L.arow; t.u_zs = L.az;
    t.w_off = L.bblk * NPAIR + y0 * L.brow; t.w_rs = L.brow; t.w_zs = L.bz;
    t.tu = site_type(x0);
    t.tw = site_type(y0);
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) t.own[a][b] = wave * NPAIR + (x0 + a) * NS3 + (y0 + b);
    return t;
  }
  t.used = lane < 51;
  const int P = t.used ? lane / 17 : 0, r = lane - P * 17;
  const PassLayout L = pass_layout(P);
  t.P = P;
  t.own[0][0] = t.own[0][1] = t.own[1][0] = t.own[1][1] = -1;
  if (r < 8) {  // O row: x = 0, y = 2r+1, 2r+2
    const int y0 = 2 * r + 1;
    t.u_off = L.ablk * NPAIR; t.u_zs = L.az;
    t.w_off = L.bblk * NPAIR + y0 * L.brow; t.w_rs = L.brow; t.w_zs = L.bz;
    t.tu = 0;
    t.tw = site_type(y0);
    t.own[0][0] = P * NPAIR + y0;
    t.own[0][1] = P * NPAIR + y0 + 1;
  } else if (r < 16) {  // O column: x = 2r-15, 2r-14, y = 0, roles swapped
    const int x0 = 2 * r - 15;
    t.u_off = L.bblk * NPAIR; t.u_zs = L.bz;
    t.w_off = L.ablk * NPAIR + x0 * L.arow; t.w_rs = L.arow; t.w_zs = L.az;
    t.tu = 0;
    t.tw = site_type(x0);
    t.own[0][0] = P * NPAIR + x0 * NS3;
    t.own[0][1] = P * NPAIR + (x0 + 1) * NS3;
  } else {  // (0, 0); the tile's second w row repeats the first and is unused
    t.u_off = L.ablk * NPAIR; t.u_zs = L.az;
    t.w_off = L.bblk * NPAIR; t.w_rs = 0; t.w_zs = L.bz;
    t.tu = t.tw = 0;
    t.own[0][0] = P * NPAIR;
  }
  t.u_rs = 0;
  return t;
}

// ====================================================================== FS3
// Distance and reciprocal of the pair stage: FP32 hardware square root and
// reciprocal in the mixed kernels; in the FP64 kernels the hardware estimates
// refined by two Newton steps (a few ulp; rsqrt_nr, rcp_nr).
template <typename T> __device__ __forceinline__ T t_dist(double r2);
template <> __device__ __forceinline__ double t_dist<double>(double r2) { return r2 * rsqrt_nr(r2); }
// the hardware square root (1 ulp); r^2 >= ~1 bohr^2 here, far from denormals
template <> __device__ __forceinline__ float t_dist<float>(double r2) { return __builtin_amdgcn_sqrtf(float(r2)); }
// (beta, gamma, r0, -) of a site pair in the kernel's precision
template <typename T> struct PairPar;
template <> struct PairPar<double> { using type = double4; };
template <> struct PairPar<float> { using type = float4; };
template <typename T> __device__ __forceinline__ T t_rcp(T x);
template <> __device__ __forceinline__ double t_rcp<double>(double x) { return rcp_nr(x); }
template <> __device__ __forceinline__ float t_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// Orbit -> dense-table entries (CSR over the 364 linear coefficients), set by
// set_orbit_tables() when a context is created.
__constant__ short c_orb_off[NLIN3 + 1];
__constant__ short c_orb_mem[NCOEF];

int set_orbit_tables(const short* off, const short* mem) {
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_orb_off), off, sizeof(short) * (NLIN3 + 1)) != hipSuccess) return -1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_orb_mem), mem, sizeof(short) * NCOEF) != hipSuccess) return -1;
  return 0;
}

// One 256-thread block per triplet:
//  pair stage:   e, e r and d ln e/dr of the 867 site pairs of a region -> LDS;
//  triple stage: 3 passes x 9 x 9 tiles of 2x2 site pairs (243 threads), each
//                accumulating G[x,y,s] = sum_tz sum_pq C M_tz over the type
//                blocks of the third molecule's sites; the tile's owner keeps
//                dE/dr of its site pairs in registers across regions and ends
//                with h = (dE/dr)/r and the site virial;
//  chain rule:   h of each site pair -> site forces; dE/dR_avg -> COM forces.
// PHI: also write the triplet's row of dE/dc (kcal/mol per unit coefficient;
// 364 SR then 364 LR) to phi_trip[t][728].
// CSM: stage the region's coefficient table in LDS; otherwise the tile threads
// read it through the vector cache (less LDS per block, more blocks per CU).
// pairpar: [2][289] (beta, gamma, r0, -) for SR and LR (FP32 copy for the mixed
// kernels); coef: [2][NCOEF] dense [tx][ty][tz][p][q][s], or with CIL in the
// interleaved layout of cil_index (read through the vector cache only).
template <typename T, bool PHI, bool CSM, bool CIL = false>
__device__ __forceinline__ void fs3_eval_body(Dev D, Consts K, const typename PairPar<T>::type* __restrict__ pairpar,
                                              const T* __restrict__ coef, double* __restrict__ phi_trip) {
  static_assert(!(CIL && (PHI || CSM)), "the interleaved table is read by the plain evaluation only");
  using T2 = typename Pair2<T>::type;
  // pair-vector basis (e, e (r - VSHIFT)); the coefficient table passed in is in the same basis
  constexpr double VSHIFT = sizeof(T) == 4 ? MIXED_VSHIFT : 0.0;
  // sR holds r (mixed: FP32-exact, from the FP32 square root, and reused by the
  // second region) or 1/r (FP64, the by-product of the distance's reciprocal
  // square root), and then h = (dE/dr)/r.
  __shared__ T sC[CSM ? NCOEF : 1];
  __shared__ double sP[3][NS3][3];
  __shared__ T2 sV[NPAIR3];
  __shared__ T sDer[NPAIR3];
  __shared__ T sR[NPAIR3];
  __shared__ double sRed[4][2];
  __shared__ double sW[2];
  __shared__ TriGeo g;  // block-uniform geometry, read with lane-varying indices
  // PHI: dE_X/dC of the dense table [tx][ty][tz][p][q][s] of the current region
  __shared__ T sDC[PHI ? NCOEF : 1];

  const long t = blockIdx.x;  // one block per triplet
  if (t >= D.cap3) return;
  const int tid = threadIdx.x;
  // The triplet's record, geometry and site positions are loaded before the
  // list length is known (t < cap3 keeps every load inside its array), so the
  // site loads wait on one level of memory latency instead of two.  A block
  // past the list may read records of an earlier call: its molecule indices
  // are clamped into range, and it returns once the length has arrived.
  static_assert(sizeof(TriGeo) % 8 == 0 && sizeof(TriGeo) / 8 <= 256, "TriGeo copy");
  const int4 tr = D.tri[t];
  double gword = 0.0, sval = 0.0, shift = 0.0;
  if (tid < (int)(sizeof(TriGeo) / 8)) gword = reinterpret_cast<const double*>(D.geo3 + t)[tid];
  if (tid < 3 * NS3 * 3) {
    // site positions in a frame centred on molecule i: s_i, s_j - x_ij and
    // s_k - x_ik, so that every site-pair vector is a plain difference
    // (x_jk = x_ik - x_ij for the derived vector of :882-884)
    const int role = tid / (NS3 * 3), rem = tid - role * NS3 * 3, c = rem - (rem / 3) * 3;
    int m = role == 0 ? tr.x : role == 1 ? tr.y : tr.z;
    m = (unsigned)m < (unsigned)D.nm ? m : 0;
    if (role > 0) shift = D.geo3[t].c[role - 1][c];
    sval = D.site[(long)m * NS3 * 3 + rem];
  }
  if (t >= D.totals[T_N3]) return;
  if (tid < (int)(sizeof(TriGeo) / 8)) reinterpret_cast<double*>(&g)[tid] = gword;
  if (tid < 3 * NS3 * 3) (&sP[0][0][0])[tid] = sval - shift;
  const int nreg = D.geo3[t].nreg, reg0 = D.geo3[t].reg0;

  T gacc[2][2] = {{T(0), T(0)}, {T(0), T(0)}};  // fd w dE/dr of the tile's site pairs
  double eacc0 = 0.0, eacc1 = 0.0;
  for (int rr = 0; rr < nreg; ++rr) {
    const int X = reg0 + rr;
    // coefficients of this region (the previous region's readers have passed
    // the barrier that ends the triple stage)
    if constexpr (CSM)
      for (int q = tid; q < NCOEF; q += 256) sC[q] = coef[X * NCOEF + q];
    if constexpr (PHI)
      for (int q = tid; q < NCOEF; q += 256) sDC[q] = T(0);
    if (rr == 0) __syncthreads();  // sP and g
    // ---------------- pair stage: e, e*r and d ln e / dr for the 867 site pairs
    // (the 17 x 17 blocks ij, ik, jk).  Thread tid < 256 takes site pair
    // q = tid of each block: one load of the pair parameters serves the three
    // blocks (they depend on the two site indices only), and both sites are
    // fixed offsets from two per-thread bases.  The last 33 pairs of each block
    // go to the first 99 threads, one each.  The mixed kernels reuse the first
    // region's distances from sR in the second.
    {
      using PP = typename PairPar<T>::type;
      auto site_pair = [&](const double* pa, const double* pb, const PP& par, int p) {
        T rt;
        if constexpr (sizeof(T) == 8) {  // FP64: recomputing costs fewer registers than keeping r
          const double vx = pa[0] - pb[0], vy = pa[1] - pb[1], vz = pa[2] - pb[2];
          const double r2 = vx * vx + vy * vy + vz * vz, ir = rsqrt_nr(r2);
          rt = r2 * ir;
          if (rr == 0) sR[p] = ir;
        } else if (rr == 0) {
          const double vx = pa[0] - pb[0], vy = pa[1] - pb[1], vz = pa[2] - pb[2];
          rt = t_dist<T>(vx * vx + vy * vy + vz * vz);
          sR[p] = rt;
        } else {
          rt = sR[p];
        }
        const T tg = t_exp<T>(-T(par.y) * (rt - T(par.z)));
        const T damp = t_rcp<T>(T(1) + tg);
        const T e = t_exp<T>(-T(par.x) * rt) * damp;
        T2 v;
        v.x = e;
        v.y = e * (rt - T(VSHIFT));
        sV[p] = v;
        sDer[p] = -T(par.x) + T(par.y) * tg * damp;
      };
      // the indices come from an opaque copy of tid: hoisted out of the region
      // loop they would hold registers through the triple stage
      int tp = tid;
      asm volatile("" : "+v"(tp));
      const double* s0 = &sP[0][0][0];
      {
        const int x = tp / NS3, y = tp - x * NS3;
        const PP par = pairpar[X * NPAIR + tp];
        const double* pa = s0 + 3 * x;
        const double* pb = s0 + 3 * y;
        // blocks ij (site x of i, site y of j), ik, jk
#pragma unroll 1
        for (int blk = 0; blk < 3; ++blk)
          site_pair(pa + (blk == 2 ? 3 * NS3 : 0), pb + (blk == 0 ? 3 * NS3 : 6 * NS3), par, blk * NPAIR + tp);
      }
      constexpr int TAIL = NPAIR - 256;
      if (tp < 3 * TAIL) {
        const int blk = tp / TAIL, q = 256 + tp - blk * TAIL;
        const int x = q / NS3, y = q - x * NS3;
        site_pair(s0 + 3 * (blk == 2 ? NS3 + x : x), s0 + 3 * ((blk == 0 ? NS3 : 2 * NS3) + y),
                  pairpar[X * NPAIR + q], blk * NPAIR + q);
      }
    }
    __syncthreads();
    // ---------------- site-triple stage.  Waves 0-2 take pass P = wave on the
    // 16 x 16 site pairs of sites 1-16, lane = one 2 x 2 tile, with P, its
    // strides and its owner indices compile-time constants; wave 3 takes the
    // O row, the O column and (O, O) of the three passes (tile_spec).  The
    // lane's tile is rebuilt each region from an opaque copy of tid: cheaper
    // than keeping it live across the region loop.
    {
      const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
      const T fdw = T(g.fd * g.w[rr]);
      const T* cbase = CSM ? sC : coef + X * NCOEF;
      if (wave < 3) {
        int lq = tid & 63;
        asm volatile("" : "+v"(lq));
        const int kx = lq >> 3, ky = lq & 7, x0 = 1 + 2 * kx, y0 = 1 + 2 * ky;
        // site_type(1 + 2k) = k + 1 for k < 2 (H, D1), else 2 + k / 2 (N1-N3)
        const int tx = kx < 2 ? kx + 1 : 2 + (kx >> 1), ty = ky < 2 ? ky + 1 : 2 + (ky >> 1);
        const int tb = (tx * NTYPE + ty) * NTYPE * 8;
        auto pass = [&](auto pc) {
          constexpr int P = decltype(pc)::value;
          constexpr PassLayout L = pass_layout(P);
          T G[2][2][2], vown[2][2][2] = {};
          T* dc = nullptr;
          if constexpr (PHI && P == 0) {
            dc = sDC + tb;
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
              for (int b2 = 0; b2 < 2; ++b2) {
                const T2 v = sV[(x0 + a) * NS3 + y0 + b2];
                vown[a][b2][0] = v.x;
                vown[a][b2][1] = v.y;
              }
          }
          contract_tile<T, 2, 2, PHI, CIL>(sV + L.ablk * NPAIR + x0 * L.arow, L.arow, L.az,
                                           sV + L.bblk * NPAIR + y0 * L.brow, L.brow, L.bz,
                                           cbase + (CIL ? 2 * (tx * NTYPE + ty) : tb), G, dc, vown);
          // the tile's site pairs (x0 + a, y0 + b2) of block P: the energy is
          // sum_xy v_xy G_xy over pass 0
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b2 = 0; b2 < 2; ++b2) {
              const int po = P * NPAIR + (x0 + a) * NS3 + y0 + b2;
              const T2 v = sV[po];
              const T ev = v.x * G[a][b2][0] + v.y * G[a][b2][1];
              gacc[a][b2] += fdw * (sDer[po] * ev + v.x * G[a][b2][1]);
              if constexpr (P == 0) {
                if (rr == 0) eacc0 += double(ev); else eacc1 += double(ev);
              }
            }
        };
        if (wave == 0) pass(std::integral_constant<int, 0>{});
        else if (wave == 1) pass(std::integral_constant<int, 1>{});
        else pass(std::integral_constant<int, 2>{});
      } else {
        int tq = tid;
        asm volatile("" : "+v"(tq));
        const TileSpec ts = tile_spec(tq);
        if (ts.used) {
          const T* cp = cbase + (CIL ? 2 * (ts.tu * NTYPE + ts.tw) : (ts.tu * NTYPE + ts.tw) * NTYPE * 8);
          T* dc = nullptr;
          if constexpr (PHI)
            if (ts.P == 0) dc = sDC + (ts.tu * NTYPE + ts.tw) * NTYPE * 8;
          T G[1][2][2], vown[1][2][2] = {};
          if constexpr (PHI)
            if (ts.P == 0)
#pragma unroll
              for (int b2 = 0; b2 < 2; ++b2) {
                const int po = ts.own[0][b2];
                const T2 v = sV[po < 0 ? 0 : po];
                vown[0][b2][0] = po < 0 ? T(0) : v.x;
                vown[0][b2][1] = po < 0 ? T(0) : v.y;
              }
          contract_tile<T, 1, 2, PHI, CIL>(sV + ts.u_off, 0, ts.u_zs, sV + ts.w_off, ts.w_rs, ts.w_zs, cp, G, dc, vown);
#pragma unroll
          for (int b2 = 0; b2 < 2; ++b2) {
            const int po = ts.own[0][b2];
            if (po < 0) continue;
            const T2 v = sV[po];
            const T ev = v.x * G[0][b2][0] + v.y * G[0][b2][1];
            gacc[0][b2] += fdw * (sDer[po] * ev + v.x * G[0][b2][1]);
            if (ts.P == 0) {
              if (rr == 0) eacc0 += double(ev); else eacc1 += double(ev);
            }
          }
        }
      }
    }
    __syncthreads();
    if constexpr (PHI) {
      if constexpr (VSHIFT != 0.0) {
        // back to the plain basis: dE/dC = (T x T x T) dE/dC', i.e. along each
        // of the three power indices of a type block, d[1] += VSHIFT d[0]
        for (int blk = tid; blk < NCOEF / 8; blk += 256) {
          T* b = sDC + blk * 8;
          for (int stride = 4; stride >= 1; stride >>= 1)
            for (int a = 0; a < 8; ++a)
              if (a & stride) b[a] += T(VSHIFT) * b[a - stride];
        }
        __syncthreads();
      }
      // orbit sums of the dense gradient, scaled to the triplet's energy in kcal/mol
      const double scale = H2KCAL * g.fd * g.w[rr];
      double* row = phi_trip + t * (2 * NLIN3);
      for (int n = tid; n < NLIN3; n += 256) {
        T a = T(0);
        for (int m = c_orb_off[n]; m < c_orb_off[n + 1]; ++m) a += sDC[c_orb_mem[m]];
        row[X * NLIN3 + n] = scale * double(a);
        if (nreg == 1) row[(1 - X) * NLIN3 + n] = 0.0;
      }
      __syncthreads();
    }
  }
  // the tile owner turns dE/dr into h = (dE/dr)/r
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  auto own_h = [&](int po, T ga) {
    if constexpr (sizeof(T) == 8) sR[po] = ga * sR[po];
    else sR[po] = ga * t_rcp<T>(sR[po]);
  };
  if (wave < 3) {
    const int lane = tid & 63, x0 = 1 + 2 * (lane >> 3), y0 = 1 + 2 * (lane & 7);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b2 = 0; b2 < 2; ++b2) own_h(wave * NPAIR + (x0 + a) * NS3 + y0 + b2, gacc[a][b2]);
  } else {
    int tq = tid;
    asm volatile("" : "+v"(tq));
    const TileSpec ts = tile_spec(tq);
    if (ts.used)
#pragma unroll
      for (int b2 = 0; b2 < 2; ++b2)
        if (ts.own[0][b2] >= 0) own_h(ts.own[0][b2], gacc[0][b2]);
  }
  // region energies (FP64, fixed reduction trees: each wavefront's DPP sum,
  // then the four wavefront sums); the barrier also publishes h.  Only waves 0
  // and 3 hold pass-0 tiles, so only they have energy terms, and only
  // switch-band triplets have a second region.
  {
    const bool has_e = wave == 0 || wave == 3;
    const double s0 = has_e ? wave_sum_lane63(eacc0) : 0.0;
    const double s1 = has_e && nreg == 2 ? wave_sum_lane63(eacc1) : 0.0;
    if ((tid & 63) == 63) {
      sRed[tid >> 6][0] = s0;
      sRed[tid >> 6][1] = s1;
    }
  }
  __syncthreads();
  const double EX0 = (sRed[0][0] + sRed[1][0]) + (sRed[2][0] + sRed[3][0]);
  const double EX1 = (sRed[0][1] + sRed[1][1]) + (sRed[2][1] + sRed[3][1]);
  const double Eb = g.w[0] * EX0 + g.w[1] * EX1;
  const double Et = g.fd * Eb;
  const double dEdR = g.dfd * Eb + g.fd * (g.wp[0] * EX0 + g.wp[1] * EX1);  // dE/dR_avg
  // ---------------- chain rule: site forces from h.  Thread q < 102 takes one
  // (molecule role, own site s, partner molecule) and sums over the partner's
  // 17 sites u.  With the positions Q of the frame above, the force on the own
  // site is -sum_u h_u (Q_own - Q_u) = T - Q_own H, H = sum_u h_u,
  // T = sum_u h_u Q_u, whichever molecule comes first in the block.  One code
  // path for every lane; the two partners of a (role, site) sit in adjacent
  // lanes and are added there, and the sums are staged in LDS for the atomics.
  // The site virial sum_pairs r dE/dr equals -sum_sites Q . F (the energy
  // depends on site differences only); each thread adds -Q_own . F of its
  // partner's share, summed over waves 0-1 with the DPP tree.
  double* sF = reinterpret_cast<double*>(sV);  // site forces [role][site][c]; the pair vectors are dead here
  static_assert(sizeof(sV) >= 3 * NS3 * 3 * sizeof(double), "site-force staging");
  double wpart = 0.0;
  if (tid < 102) {
    const int role = tid / 34, r = tid - role * 34, s = r >> 1, half = r & 1;
    const int blk = half ? (role == 0 ? 1 : 2) : (role == 2 ? 1 : 0);  // ij, ik, jk
    const int pm = half ? (role == 2 ? 1 : 2) : (role == 0 ? 1 : 0);   // partner role
    const bool first = role == 0 || (role == 1 && half);
    const int hstride = first ? 1 : NS3;
    const T* hp = sR + blk * NPAIR + (first ? s * NS3 : s);
    const double* pp = pm == 0 ? &sP[0][0][0] : pm == 1 ? &sP[1][0][0] : &sP[2][0][0];
    double H = 0.0, T0 = 0.0, T1 = 0.0, T2 = 0.0;
#pragma unroll
    for (int u = 0; u < NS3; ++u) {
      const double h = double(hp[u * hstride]);
      H += h;
      T0 = fma(h, pp[3 * u], T0);
      T1 = fma(h, pp[3 * u + 1], T1);
      T2 = fma(h, pp[3 * u + 2], T2);
    }
    double F0 = fma(-sP[role][s][0], H, T0);
    double F1 = fma(-sP[role][s][1], H, T1);
    double F2 = fma(-sP[role][s][2], H, T2);
    wpart = -(sP[role][s][0] * F0 + sP[role][s][1] * F1 + sP[role][s][2] * F2);
    F0 += __shfl_xor(F0, 1, 64);
    F1 += __shfl_xor(F1, 1, 64);
    F2 += __shfl_xor(F2, 1, 64);
    if (!half) {
      sF[role * NS3 * 3 + s * 3] = F0;
      sF[role * NS3 * 3 + s * 3 + 1] = F1;
      sF[role * NS3 * 3 + s * 3 + 2] = F2;
    }
  } else if (tid >= 192 && tid < 201) {
    // COM forces from the R_avg dependence of f_d and w (:1130-1166, 1613-1665)
    const int q = tid - 192, role = q / 3, c = q - role * 3;
    const double uij = g.c[0][c] * rcp_nr(g.R[0]), uik = g.c[1][c] * rcp_nr(g.R[1]);
    const double ujk = g.c[2][c] * rcp_nr(g.R[2]);
    const double g3 = -dEdR / 3.0;
    double F;
    if (role == 0) F = g3 * (uij + uik);
    else if (role == 1) F = g3 * (-uij + ujk);
    else F = g3 * (-uik - ujk);
    fx_add(D, (long)g.m[role] * ACC_W + ACC_C3 + c, F);
  }
  if (wave < 2) {
    const double ws = wave_sum_lane63(wpart);
    if ((tid & 63) == 63) sW[wave] = ws;
  }
  __syncthreads();
  // one fixed-point atomic per site-force component, consecutive lanes on
  // consecutive accumulator entries (51 per molecule)
  if (tid < 3 * NS3 * 3) {
    const int role = tid / (NS3 * 3);
    fx_add(D, (long)g.m[role] * ACC_W + (tid - role * NS3 * 3), sF[tid]);
  }
  if (tid == 0) {
    D.ew3[2 * t] = Et;
    D.ew3[2 * t + 1] = (sW[0] + sW[1]) + g.Ravg * dEdR;
  }
}

// Accumulators (hartree/bohr) -> the caller's arrays in 10 J/mol/A: site
// forces added (:574-585, 1964-1977), COM forces overwritten.  Thread per
// (molecule, accumulator) t.
__device__ __forceinline__ void finish_one(const Dev& D, long t, int ld, int nsite, int do2, int do3, double* fxs,
                                           double* fys, double* fzs, double* fcom2, double* fcom3) {
  if (t >= (long)D.nm * ACC_W) return;
  const int m = (int)(t / ACC_W), a = (int)(t - (long)m * ACC_W);
  const double conv = H2KCAL * 418.4 / BOHR2A;
  const double v = fx_decode(D.acc + 2 * t) * conv;
  if (a < NS3 * 3) {
    const int s = a / 3, c = a - s * 3;
    if (s < nsite) {
      double* f = c == 0 ? fxs : c == 1 ? fys : fzs;
      f[(long)ld * m + s] += v;
    }
  } else if (a < ACC_C3) {
    if (do2) fcom2[(long)(a - ACC_C2) * D.nm + m] = v;
  } else if (do3) {
    fcom3[(long)(a - ACC_C3) * D.nm + m] = v;
  }
}

// Energy/virial reduction, level 1: one 256-thread block per box-aligned chunk
// q of RED_CHUNK entries (per-triplet FS3 or per-(spectator, part) FS2
// values), fixed tree.  Called by a whole block.
__device__ __forceinline__ void energy_chunk(const Dev& D, int which, long q) {
  __shared__ double red[2][4];
  if (q >= (which == 3 ? D.box_c3 : D.box_c2)[D.nbox]) return;
  const int* bt = which == 3 ? D.box_t3 : D.box_t2;
  const int* bc = which == 3 ? D.box_c3 : D.box_c2;
  const double* ew = which == 3 ? D.ew3 : D.ew2;
  double* out = which == 3 ? D.red3 : D.red2;
  int lo = 0, hi = D.nbox;  // box owning chunk q: bc[b] <= q < bc[b+1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (bc[mid] <= q) lo = mid; else hi = mid;
  }
  const int b = lo;
  const long t0 = bt[b] + (q - bc[b]) * (long)RED_CHUNK;
  const long t1 = t0 + RED_CHUNK < bt[b + 1] ? t0 + RED_CHUNK : bt[b + 1];
  double e = 0.0, w = 0.0;
  for (long t = t0 + threadIdx.x; t < t1; t += 256) {
    e += ew[2 * t];
    w += ew[2 * t + 1];
  }
  e = wave_sum(e);
  w = wave_sum(w);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = e;
    red[1][threadIdx.x >> 6] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[2 * q] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    out[2 * q + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// Level 2 of the energy sums, one wavefront per box: lane l sums chunks l,
// l+64, ... of its box in order, then a fixed butterfly -> {qfs2, qfs3
// (kcal/mol), vir increment (10 J/mol)}.  The order depends only on the chunk
// layout, not on timing.
__device__ __forceinline__ void box_energies(const Dev& D, int b, int do2, int do3, double* out) {
  const int lane = threadIdx.x & 63;
  double e2 = 0.0, w2 = 0.0, e3 = 0.0, w3 = 0.0;
  if (do2)
    for (int q = D.box_c2[b] + lane; q < D.box_c2[b + 1]; q += 64) {
      e2 += D.red2[2 * q];
      w2 += D.red2[2 * q + 1];
    }
  if (do3)
    for (int q = D.box_c3[b] + lane; q < D.box_c3[b + 1]; q += 64) {
      e3 += D.red3[2 * q];
      w3 += D.red3[2 * q + 1];
    }
  e2 = wave_sum(e2);
  w2 = wave_sum(w2);
  e3 = wave_sum(e3);
  w3 = wave_sum(w3);
  if (lane == 0) {
    out[3 * b + 0] = e2 * H2KCAL;
    out[3 * b + 1] = H2KCAL * e3;
    out[3 * b + 2] = w2 * H2KCAL * 418.4 + w3 * H2KCAL * 418.4;
  }
}

// The call's epilogue in one launch: blocks [0, nfin) write the accumulators
// back, the next nch3 blocks reduce FS3 energy chunks and the last nch2 blocks
// FS2 energy chunks.  The last block to finish (a counter in totals) then forms
// the per-box sums and, with host_totals, copies the call's counters to the
// host (pinned memory mapped into the device's address space).  Nothing but
// the counters is written when a list overflowed.
__global__ __launch_bounds__(256) void k_epilogue(Dev D, int ld, int nsite, int do2, int do3, double* fxs,
                                                  double* fys, double* fzs, double* fcom2, double* fcom3,
                                                  long nfin, long nch3, long nch2, double* energies,
                                                  int* host_totals) {
  __shared__ bool last;
  const bool over = lists_overflowed(D);
  const long b = blockIdx.x;
  if (!over) {
    if (b < nfin)
      finish_one(D, b * blockDim.x + threadIdx.x, ld, nsite, do2, do3, fxs, fys, fzs, fcom2, fcom3);
    else if (b < nfin + nch3)
      energy_chunk(D, 3, b - nfin);
    else if (b < nfin + nch3 + nch2)
      energy_chunk(D, 2, b - nfin - nch3);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();  // this block's results before its count
    last = atomicAdd(D.totals + T_DONE, 1) == (int)gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();  // every other block's results after their counts
  if (!over)
    for (int bx = threadIdx.x >> 6; bx < D.nbox; bx += 4) box_energies(D, bx, do2, do3, energies);
  if (host_totals && threadIdx.x < T_COUNT) {
    host_totals[threadIdx.x] = __atomic_load_n(D.totals + threadIdx.x, __ATOMIC_RELAXED);
    __threadfence_system();
  }
}

// Coefficients come through the vector cache; LDS per block is ~29 KB (FP64,
// 94 VGPRs, 96 allocated: 5 blocks per CU) and ~15 KB (mixed, 64 VGPRs: 8
// blocks per CU), neither with spilled registers.  The mixed kernels take the table in the
// shifted basis (MIXED_VSHIFT), the FP64 kernels the plain one; k_fs3_eval_f64
// takes it in the interleaved layout (cil_index).
__global__ __launch_bounds__(256, 5) void k_fs3_eval_f64(Dev D, Consts K, const double4* __restrict__ pairpar,
                                                         const double* __restrict__ coef_il) {
  fs3_eval_body<double, false, false, true>(D, K, pairpar, coef_il, nullptr);
}

__global__ __launch_bounds__(256, 8) void k_fs3_eval_f32(Dev D, Consts K, const float4* __restrict__ pairpar,
                                                         const float* __restrict__ coef) {
  fs3_eval_body<float, false, false>(D, K, pairpar, coef, nullptr);
}

// The same with the region's coefficient table staged in LDS (a tuning choice).
__global__ __launch_bounds__(256, 3) void k_fs3_eval_f64_lds(Dev D, Consts K, const double4* __restrict__ pairpar,
                                                             const double* __restrict__ coef) {
  fs3_eval_body<double, false, true>(D, K, pairpar, coef, nullptr);
}

__global__ __launch_bounds__(256, 7) void k_fs3_eval_f32_lds(Dev D, Consts K, const float4* __restrict__ pairpar,
                                                             const float* __restrict__ coef) {
  fs3_eval_body<float, false, true>(D, K, pairpar, coef, nullptr);
}

// The same evaluation plus the per-triplet coefficient-gradient rows.
__global__ __launch_bounds__(256, 2) void k_fs3_phi_f64(Dev D, Consts K, const double4* __restrict__ pairpar,
                                                        const double* __restrict__ coef, double* phi_trip) {
  fs3_eval_body<double, true, true>(D, K, pairpar, coef, phi_trip);
}

__global__ __launch_bounds__(256, 4) void k_fs3_phi_f32(Dev D, Consts K, const float4* __restrict__ pairpar,
                                                        const float* __restrict__ coef, double* phi_trip) {
  fs3_eval_body<float, true, true>(D, K, pairpar, coef, phi_trip);
}

// Coefficient gradient of each box: the sum, in triplet order, of its
// triplets' rows (boxes own contiguous triplet ranges).  Thread per (box, coefficient).
__global__ __launch_bounds__(256) void k_phi_boxes(Dev D, const double* __restrict__ phi_trip, double* phi) {
  const int b = blockIdx.y;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= D.nbox || n >= 2 * NLIN3 || lists_overflowed(D)) return;
  double a = 0.0;
  for (int t = D.box_t3[b]; t < D.box_t3[b + 1]; ++t) a += phi_trip[(long)t * 2 * NLIN3 + n];
  phi[(long)b * 2 * NLIN3 + n] = a;
}

}  // namespace ccpol3b
