// FS3 and FS2 evaluation kernels, deterministic per-molecule gathers and the
// per-box energy/virial reductions.
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
#include "ccpol3b_internal.h"

namespace ccpol3b {

template <typename T> struct Pair2;
template <> struct Pair2<double> { using type = double2; };
template <> struct Pair2<float> { using type = float2; };

__device__ __forceinline__ double minimg_e(double d, double L) {
  return L > 0.0 ? d - L * round(d / L) : d;
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

template <typename T> __device__ __forceinline__ T t_exp(T x);
template <> __device__ __forceinline__ double t_exp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float t_exp<float>(float x) { return __expf(x); }

// 2x2 site tiles of one molecule's 17 sites, aligned to the type blocks
// (O | H H | D1 D1 | N1 N1 | N1 N1 | N2 N2 | N2 N2 | N3 N3 | N3 N3); -1 pads O.
__constant__ int c_tile0[9] = {0, 1, 3, 5, 7, 9, 11, 13, 15};
__constant__ int c_tile1[9] = {-1, 2, 4, 6, 8, 10, 12, 14, 16};
__constant__ int c_tilet[9] = {0, 1, 2, 3, 3, 4, 4, 5, 5};
__constant__ int c_z0[6] = {0, 1, 3, 5, 9, 13};
__constant__ int c_z1[6] = {1, 3, 5, 9, 13, 17};

// ====================================================================== FS3
// One 256-thread block per triplet (grid-stride).  pairpar: [2][289] (beta,
// gamma, r0, -) for SR and LR; coef: [2][NCOEF] dense [tx][ty][tz][p][q][s].
template <typename T>
__global__ __launch_bounds__(256, 2) void k_fs3_eval(Dev D, Consts K,
                                                      const double4* __restrict__ pairpar,
                                                      const T* __restrict__ coef) {
  using T2 = typename Pair2<T>::type;
  __shared__ T sC[2 * NCOEF];
  __shared__ double sP[3][NS3][3];
  __shared__ T2 sV[NPAIR3];
  __shared__ T sDer[NPAIR3];
  __shared__ double sR[NPAIR3];
  __shared__ T sG[NPAIR3];
  __shared__ double sRed[4];

  const int tid = threadIdx.x;
  for (int q = tid; q < 2 * NCOEF; q += 256) sC[q] = coef[q];

  const long n3 = D.totals[0];
  for (long t = blockIdx.x; t < n3; t += gridDim.x) {
    const int4 tr = D.tri[t];
    const int mi = tr.x, mj = tr.y, mk = tr.z;
    const double L = D.boxL[D.mol_box[mi]];
    // COM vectors as the reference forms them (:862-884)
    double cij[3], cik[3], cjk[3];
    {
      const double pi_[3] = {D.cx[mi], D.cy[mi], D.cz[mi]};
      const double pj_[3] = {D.cx[mj], D.cy[mj], D.cz[mj]};
      const double pk_[3] = {D.cx[mk], D.cy[mk], D.cz[mk]};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double dij = pi_[c] - pj_[c], dik = pi_[c] - pk_[c], djk = pj_[c] - pk_[c];
        if (L > 0.0) {
          const double nij = round(dij / L), nik = round(dik / L);
          cij[c] = dij - L * nij;
          cik[c] = dik - L * nik;
          cjk[c] = djk + L * nij - L * nik;
        } else {
          cij[c] = dij;
          cik[c] = dik;
          cjk[c] = djk;
        }
      }
    }
    const double Rij = sqrt(cij[0] * cij[0] + cij[1] * cij[1] + cij[2] * cij[2]);
    const double Rik = sqrt(cik[0] * cik[0] + cik[1] * cik[1] + cik[2] * cik[2]);
    const double Rjk = sqrt(cjk[0] * cjk[0] + cjk[1] * cjk[1] + cjk[2] * cjk[2]);
    const double Ravg = (Rij + Rjk + Rik) / 3.0;
    const double pom = exp(K.alp3 * (Ravg - K.r0dmp3));
    const double pp = 1.0 / (1.0 + pom);
    const double fd = pow(pp, K.delta33);
    const double dfd = -K.alp3 * K.delta33 * pom * pp * fd;  // d f_d / d R_avg
    int nreg, reg0;
    double w0 = 1.0, wp0 = 0.0, w1 = 0.0, wp1 = 0.0;
    if (Ravg <= K.rmin) {
      nreg = 1; reg0 = 0;
    } else if (Ravg >= K.rmax) {
      nreg = 1; reg0 = 1;
    } else {
      const double span = K.rmax - K.rmin;
      const double c = (Ravg - K.rmin) / span;
      const double sm = 1.0 + 2.0 * c * c * c - 3.0 * c * c;
      const double dsm = (6.0 * c * c - 6.0 * c) / span;
      nreg = 2; reg0 = 0;
      w0 = sm; wp0 = dsm;
      w1 = 1.0 - sm; wp1 = -dsm;
    }
    if (tid == 0) atomicAdd(D.totals + 4 + (nreg == 2 ? 1 : 2 * reg0), 1);

    if (tid < 3 * NS3 * 3) {
      const int role = tid / (NS3 * 3), rem = tid - role * NS3 * 3;
      const int m = role == 0 ? mi : role == 1 ? mj : mk;
      (&sP[0][0][0])[tid] = D.site[(long)m * NS3 * 3 + rem];
    }
    for (int p = tid; p < NPAIR3; p += 256) sG[p] = T(0);
    __syncthreads();

    double eacc0 = 0.0, eacc1 = 0.0;
    for (int rr = 0; rr < nreg; ++rr) {
      const int X = reg0 + rr;
      // ---------------- pair stage: e, e*r and d ln e / dr for the 867 site pairs
      for (int p = tid; p < NPAIR3; p += 256) {
        const int blk = p / NPAIR, rem = p - blk * NPAIR;
        const int x = rem / NS3, y = rem - x * NS3;
        const int ra = blk == 2 ? 1 : 0, rb = blk == 0 ? 1 : 2;
        const double cvx = blk == 0 ? cij[0] : blk == 1 ? cik[0] : cjk[0];
        const double cvy = blk == 0 ? cij[1] : blk == 1 ? cik[1] : cjk[1];
        const double cvz = blk == 0 ? cij[2] : blk == 1 ? cik[2] : cjk[2];
        const double vx = cvx + (sP[ra][x][0] - sP[rb][y][0]);
        const double vy = cvy + (sP[ra][x][1] - sP[rb][y][1]);
        const double vz = cvz + (sP[ra][x][2] - sP[rb][y][2]);
        const double r = sqrt(vx * vx + vy * vy + vz * vz);
        const double4 par = pairpar[X * NPAIR + rem];
        const T rt = T(r);
        const T tg = t_exp<T>(-T(par.y) * (rt - T(par.z)));
        const T damp = T(1) / (T(1) + tg);
        const T e = t_exp<T>(-T(par.x) * rt) * damp;
        T2 v;
        v.x = e;
        v.y = e * rt;
        sV[p] = v;
        sDer[p] = -T(par.x) + T(par.y) * tg * damp;
        if (rr == 0) sR[p] = r;
      }
      __syncthreads();
      // ---------------- site-triple stage: 3 passes x 9 x 9 tiles of 2x2 site pairs
      if (tid < 243) {
        const int P = tid / 81, rem = tid - P * 81;
        const int xb = rem / 9, yb = rem - xb * 9;
        const int x0 = c_tile0[xb], x1r = c_tile1[xb], tx = c_tilet[xb];
        const int y0 = c_tile0[yb], y1r = c_tile1[yb], ty = c_tilet[yb];
        const int x1 = x1r < 0 ? x0 : x1r, y1 = y1r < 0 ? y0 : y1r;
        // pair-array addressing of v_xz and v_yz for this pass
        int xzb, xzs, xzz, yzb, yzs, yzz, outb;
        if (P == 0) {        // X=i, Y=j, Z=k
          xzb = NPAIR; xzs = NS3; xzz = 1; yzb = 2 * NPAIR; yzs = NS3; yzz = 1; outb = 0;
        } else if (P == 1) { // X=i, Y=k, Z=j
          xzb = 0; xzs = NS3; xzz = 1; yzb = 2 * NPAIR; yzs = 1; yzz = NS3; outb = NPAIR;
        } else {             // X=j, Y=k, Z=i
          xzb = 0; xzs = 1; xzz = NS3; yzb = NPAIR; yzs = 1; yzz = NS3; outb = 2 * NPAIR;
        }
        T G[2][2][2];
#pragma unroll
        for (int a = 0; a < 8; ++a) (&G[0][0][0])[a] = T(0);
        const T* Cx = sC + X * NCOEF + (tx * NTYPE + ty) * NTYPE * 8;
#pragma unroll 1
        for (int tz = 0; tz < NTYPE; ++tz) {
          T M[2][2][2][2];
#pragma unroll
          for (int a = 0; a < 16; ++a) (&M[0][0][0][0])[a] = T(0);
#pragma unroll 1
          for (int z = c_z0[tz]; z < c_z1[tz]; ++z) {
            const T2 a0 = sV[xzb + x0 * xzs + z * xzz];
            const T2 a1 = sV[xzb + x1 * xzs + z * xzz];
            const T2 b0 = sV[yzb + y0 * yzs + z * yzz];
            const T2 b1 = sV[yzb + y1 * yzs + z * yzz];
            const T av[2][2] = {{a0.x, a0.y}, {a1.x, a1.y}};
            const T bv[2][2] = {{b0.x, b0.y}, {b1.x, b1.y}};
#pragma unroll
            for (int xs = 0; xs < 2; ++xs)
#pragma unroll
              for (int ys = 0; ys < 2; ++ys)
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                  for (int q = 0; q < 2; ++q) M[xs][ys][p][q] += bv[ys][p] * av[xs][q];
          }
          const T* c8 = Cx + tz * 8;
          T cc[8];
#pragma unroll
          for (int a = 0; a < 8; ++a) cc[a] = c8[a];
#pragma unroll
          for (int xs = 0; xs < 2; ++xs)
#pragma unroll
            for (int ys = 0; ys < 2; ++ys)
#pragma unroll
              for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int q = 0; q < 2; ++q)
#pragma unroll
                  for (int s = 0; s < 2; ++s)
                    G[xs][ys][s] += cc[(p * 2 + q) * 2 + s] * M[xs][ys][p][q];
        }
        const T fdw = T(fd * (rr == 0 ? w0 : w1));
#pragma unroll
        for (int xs = 0; xs < 2; ++xs) {
          if (xs == 1 && x1r < 0) continue;
          const int x = xs ? x1 : x0;
#pragma unroll
          for (int ys = 0; ys < 2; ++ys) {
            if (ys == 1 && y1r < 0) continue;
            const int y = ys ? y1 : y0;
            const int po = outb + x * NS3 + y;
            const T2 v = sV[po];
            const T der = sDer[po];
            const T ev = v.x * G[xs][ys][0] + v.y * G[xs][ys][1];
            const T dEdr = der * ev + v.x * G[xs][ys][1];
            sG[po] += fdw * dEdr;
            if (P == 0) {
              if (rr == 0) eacc0 += double(ev); else eacc1 += double(ev);
            }
          }
        }
      }
      __syncthreads();
    }
    // region energies (FP64 reduction)
    const double EX0 = block_sum256(eacc0, sRed);
    const double EX1 = nreg == 2 ? block_sum256(eacc1, sRed) : 0.0;
    const double Eb = w0 * EX0 + w1 * EX1;
    double dEdR = wp0 * EX0 + wp1 * EX1;
    const double Et = fd * Eb;
    dEdR = dfd * Eb + fd * dEdR;  // dE/dR_avg
    // ---------------- chain rule: h = (dE/dr)/r; site virial sum r dE/dr
    double wsum = 0.0;
    for (int p = tid; p < NPAIR3; p += 256) {
      const double g = double(sG[p]);
      const double r = sR[p];
      wsum += g * r;
      sR[p] = g / r;
    }
    const double Wsites = block_sum256(wsum, sRed);  // includes the barrier before reuse of sR
    double* slot = D.slot3 + t * (3 * SLOT3);
    if (tid < 3 * NS3 * 3) {
      const int role = tid / (NS3 * 3), rem = tid - role * NS3 * 3;
      const int s = rem / 3, c = rem - s * 3;
      const double aij = c == 0 ? cij[0] : c == 1 ? cij[1] : cij[2];
      const double aik = c == 0 ? cik[0] : c == 1 ? cik[1] : cik[2];
      const double ajk = c == 0 ? cjk[0] : c == 1 ? cjk[1] : cjk[2];
      double F = 0.0;
      if (role == 0) {
        for (int y = 0; y < NS3; ++y)
          F -= sR[s * NS3 + y] * (aij + (sP[0][s][c] - sP[1][y][c]));
        for (int z = 0; z < NS3; ++z)
          F -= sR[NPAIR + s * NS3 + z] * (aik + (sP[0][s][c] - sP[2][z][c]));
      } else if (role == 1) {
        for (int x = 0; x < NS3; ++x)
          F += sR[x * NS3 + s] * (aij + (sP[0][x][c] - sP[1][s][c]));
        for (int z = 0; z < NS3; ++z)
          F -= sR[2 * NPAIR + s * NS3 + z] * (ajk + (sP[1][s][c] - sP[2][z][c]));
      } else {
        for (int x = 0; x < NS3; ++x)
          F += sR[NPAIR + x * NS3 + s] * (aik + (sP[0][x][c] - sP[2][s][c]));
        for (int y = 0; y < NS3; ++y)
          F += sR[2 * NPAIR + y * NS3 + s] * (ajk + (sP[1][y][c] - sP[2][s][c]));
      }
      slot[role * SLOT3 + s * 3 + c] = F;
    } else if (tid < 3 * NS3 * 3 + 9) {
      // COM forces from the R_avg dependence of f_d and w (:1130-1166, 1613-1665)
      const int q = tid - 3 * NS3 * 3, role = q / 3, c = q - role * 3;
      const double aij = c == 0 ? cij[0] : c == 1 ? cij[1] : cij[2];
      const double aik = c == 0 ? cik[0] : c == 1 ? cik[1] : cik[2];
      const double ajk = c == 0 ? cjk[0] : c == 1 ? cjk[1] : cjk[2];
      const double g3 = -dEdR / 3.0;
      double F;
      if (role == 0) F = g3 * (aij / Rij + aik / Rik);
      else if (role == 1) F = g3 * (-aij / Rij + ajk / Rjk);
      else F = g3 * (-aik / Rik - ajk / Rjk);
      slot[role * SLOT3 + NS3 * 3 + c] = F;
    }
    if (tid == 0) {
      D.ew3[2 * t] = Et;
      D.ew3[2 * t + 1] = Wsites + Ravg * dEdR;
    }
    __syncthreads();
  }
}

// FS3 gather: one wavefront per molecule m.  Its triplets are those of its own
// segment (role i) and, for each neighbour i < m, the entries of segment i
// with j == m or k == m; they are summed in ascending (i, t) order so the
// result is bitwise reproducible.  Adds to the site forces, sets COM forces.
__global__ __launch_bounds__(256) void k_fs3_gather(Dev D, int ld, double* fxs, double* fys,
                                                     double* fzs, double* fcom) {
  const int m = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (m >= D.nm) return;
  const int* nb = D.nb3 + (long)m * D.cap_nb;
  const int n = D.nb3_cnt[m];
  int u0 = 0;
  while (u0 < n && nb[u0] < m) ++u0;
  double acc = 0.0;
  for (int a = 0; a <= u0; ++a) {
    const int i = a < u0 ? nb[a] : m;
    const int s0 = D.off3[i], s1 = D.off3[i + 1];
    for (int base = s0; base < s1; base += 64) {
      const int t = base + lane;
      int role = -1;
      if (t < s1) {
        if (i == m) role = 0;
        else {
          const int4 q = D.tri[t];
          role = q.y == m ? 1 : q.z == m ? 2 : -1;
        }
      }
      uint64_t mask = __ballot(role >= 0);
      while (mask) {
        const int b = __ffsll((unsigned long long)mask) - 1;
        mask &= mask - 1;
        const int rb = __shfl(role, b, 64);
        if (lane < SLOT3) acc += D.slot3[(long)(base + b) * 3 * SLOT3 + rb * SLOT3 + lane];
      }
    }
  }
  const double conv = H2KCAL * 418.4 / BOHR2A;
  if (lane < NS3 * 3) {
    const int s = lane / 3, c = lane - s * 3;
    double* f = c == 0 ? fxs : c == 1 ? fys : fzs;
    f[(long)ld * m + s] += acc * conv;
  } else if (lane < SLOT3) {
    const int c = lane - NS3 * 3;
    fcom[(long)c * D.nm + m] = acc * conv;
  }
}

// ====================================================================== FS2
// FS2 (threebody_potential.f90:101-674).  For an accepted item (k; i<j):
//   E = f(R_ki) f(R_kj) S_ij T_k,  S_ij = eta sum_{a,b in O,H,H} exp(-beta r_ab),
//   T_k = sum_{ks=1..5} q_ks (1/|r_ks-R_i| + 1/|r_ks-R_j| - 1/|r_ks-R_i^+| - 1/|r_ks-R_j^+|)
// with R_i^+ = R_i + g u_ij, R_j^+ = R_j - g u_ij and f the damping of :339-349.
// The pair pass owns everything that belongs to i and j (S_ij and its site
// gradient, COM forces of i and j); the spectator pass owns everything that
// belongs to k.  Both recompute T_k, so neither needs cross-thread scatter.
__constant__ double c_q2[NS2] = {0.258504607653328389, 0.564050362569283648, 0.564050362569283648,
                                 -0.693302666395947953, -0.693302666395947953};
constexpr double ETA3 = 0.48056977605478400;
constexpr double BETA3 = 0.85836606579445000;
constexpr double G3 = 0.400000000000000;

struct Fs2Geo {
  double xij[3], xki[3], xkj[3];
  double Rij, Rki, Rkj;
};

__device__ __forceinline__ void fs2_damp(double R, const Consts& K, double& f, double& fp) {
  const double pom = exp(K.alp3 * (R - K.r0dmp2));
  const double pp = 1.0 / (1.0 + pom);
  f = pow(pp, K.delta33);
  fp = -K.alp3 * K.delta33 * pom * pp * f;  // df/dR
}

// T_k and its derivatives.  sk: spectator site offsets [5][3].
// dTdsk: dT/d(site ks), dTdRi, dTdRj: explicit COM derivatives, Tl: dT/dlambda
// under uniform scaling of all vectors except the fixed shift g.
template <bool WANT_SITES, bool WANT_COM>
__device__ __forceinline__ double fs2_T(const Fs2Geo& G, const double (*sk)[3], double (*dTdsk)[3],
                                        double* dTdRi, double* dTdRj, double& Tl) {
  const double uij[3] = {G.xij[0] / G.Rij, G.xij[1] / G.Rij, G.xij[2] / G.Rij};
  double T = 0.0;
  Tl = 0.0;
  if (WANT_COM)
    for (int c = 0; c < 3; ++c) dTdRi[c] = dTdRj[c] = 0.0;
  for (int ks = 0; ks < NS2; ++ks) {
    const double q = c_q2[ks];
    double dci[3], dcj[3], dqi[3], dqj[3];
    for (int c = 0; c < 3; ++c) {
      dci[c] = G.xki[c] + sk[ks][c];
      dcj[c] = G.xkj[c] + sk[ks][c];
      dqi[c] = dci[c] - G3 * uij[c];
      dqj[c] = dcj[c] + G3 * uij[c];
    }
    const double ici = 1.0 / sqrt(dci[0] * dci[0] + dci[1] * dci[1] + dci[2] * dci[2]);
    const double icj = 1.0 / sqrt(dcj[0] * dcj[0] + dcj[1] * dcj[1] + dcj[2] * dcj[2]);
    const double iqi = 1.0 / sqrt(dqi[0] * dqi[0] + dqi[1] * dqi[1] + dqi[2] * dqi[2]);
    const double iqj = 1.0 / sqrt(dqj[0] * dqj[0] + dqj[1] * dqj[1] + dqj[2] * dqj[2]);
    const double ici3 = ici * ici * ici, icj3 = icj * icj * icj;
    const double iqi3 = iqi * iqi * iqi, iqj3 = iqj * iqj * iqj;
    const double term0 = ici + icj - iqi - iqj;
    T += q * term0;
    const double pqi = dqi[0] * uij[0] + dqi[1] * uij[1] + dqi[2] * uij[2];
    const double pqj = dqj[0] * uij[0] + dqj[1] * uij[1] + dqj[2] * uij[2];
    Tl += q * (-term0 + G3 * pqi * iqi3 - G3 * pqj * iqj3);
    if (WANT_SITES)
      for (int c = 0; c < 3; ++c)
        dTdsk[ks][c] = q * (-dci[c] * ici3 - dcj[c] * icj3 + dqi[c] * iqi3 + dqj[c] * iqj3);
    if (WANT_COM) {
      // projector P = I - u u^T applied to dqi, dqj
      for (int c = 0; c < 3; ++c) {
        const double Pqi = dqi[c] - pqi * uij[c];
        const double Pqj = dqj[c] - pqj * uij[c];
        dTdRi[c] += q * (dci[c] * ici3 - (dqi[c] + G3 * Pqi / G.Rij) * iqi3 + G3 * Pqj / G.Rij * iqj3);
        dTdRj[c] += q * (dcj[c] * icj3 + G3 * Pqi / G.Rij * iqi3 - (dqj[c] + G3 * Pqj / G.Rij) * iqj3);
      }
    }
  }
  return T;
}

__device__ __forceinline__ void fs2_geometry(const Dev& D, int k, int i, int j, Fs2Geo& G) {
  const double L = D.boxL[D.mol_box[i]];
  const double pi_[3] = {D.cx[i], D.cy[i], D.cz[i]};
  const double pj_[3] = {D.cx[j], D.cy[j], D.cz[j]};
  const double pk_[3] = {D.cx[k], D.cy[k], D.cz[k]};
  for (int c = 0; c < 3; ++c) {
    G.xij[c] = minimg_e(pi_[c] - pj_[c], L);
    G.xki[c] = minimg_e(pk_[c] - pi_[c], L);
    G.xkj[c] = G.xki[c] + G.xij[c];  // same periodic image (:325-328)
  }
  G.Rij = sqrt(G.xij[0] * G.xij[0] + G.xij[1] * G.xij[1] + G.xij[2] * G.xij[2]);
  G.Rki = sqrt(G.xki[0] * G.xki[0] + G.xki[1] * G.xki[1] + G.xki[2] * G.xki[2]);
  G.Rkj = sqrt(G.xkj[0] * G.xkj[0] + G.xkj[1] * G.xkj[1] + G.xkj[2] * G.xkj[2]);
}

// Pair pass: one wavefront per FS2 pair (i<j).  Lanes 0..8 build S_ij over the
// 3x3 O/H site pairs; then lanes run over the spectator candidates k in nb2(i).
// Slot layout: [S, S_lambda, F_i sites 0..2 (9), F_COM_i (3), F_j sites (9), F_COM_j (3)].
__global__ __launch_bounds__(256) void k_fs2_pair(Dev D, Consts K) {
  const long p = (long)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  const long np = D.totals[1];
  if (p >= np) return;
  const int2 pr = D.pairs2[p];
  const int i = pr.x, j = pr.y;
  const double L = D.boxL[D.mol_box[i]];
  double xij[3];
  xij[0] = minimg_e(D.cx[i] - D.cx[j], L);
  xij[1] = minimg_e(D.cy[i] - D.cy[j], L);
  xij[2] = minimg_e(D.cz[i] - D.cz[j], L);
  // S_ij and dS/d(site) over the 9 site pairs
  double e = 0.0, el = 0.0, gv[3] = {0.0, 0.0, 0.0};
  if (lane < 9) {
    const int is = lane / 3, js = lane - is * 3;
    const double* si = D.site + ((long)i * NS3 + is) * 3;
    const double* sj = D.site + ((long)j * NS3 + js) * 3;
    double v[3];
    for (int c = 0; c < 3; ++c) v[c] = xij[c] + (si[c] - sj[c]);
    const double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    e = ETA3 * exp(-BETA3 * r);
    el = -BETA3 * r * e;
    for (int c = 0; c < 3; ++c) gv[c] = -BETA3 * e * v[c] / r;  // dS/dv
  }
  const double S = wave_sum(e);
  const double Sl = wave_sum(el);
  // dS/d(site is of i) = sum_js gv ; dS/d(site js of j) = -sum_is gv  (lanes 0..8)
  double dSi[3], dSj[3];
  {
    const int is = lane / 3, js = lane - (lane / 3) * 3;
    for (int c = 0; c < 3; ++c) {
      const double g0 = __shfl(gv[c], 3 * (lane % 3) + 0, 64);
      const double g1 = __shfl(gv[c], 3 * (lane % 3) + 1, 64);
      const double g2 = __shfl(gv[c], 3 * (lane % 3) + 2, 64);
      dSi[c] = (g0 + g1) + g2;  // lane%3 plays the role of is
      const double h0 = __shfl(gv[c], 0 * 3 + (lane % 3), 64);
      const double h1 = __shfl(gv[c], 1 * 3 + (lane % 3), 64);
      const double h2 = __shfl(gv[c], 2 * 3 + (lane % 3), 64);
      dSj[c] = -((h0 + h1) + h2);  // lane%3 plays the role of js
    }
    (void)is; (void)js;
  }
  // spectators
  const int* nbi = D.nb2 + (long)i * D.cap_nb;
  const int ni = D.nb2_cnt[i];
  double Fi[3] = {0, 0, 0}, Fj[3] = {0, 0, 0}, WT = 0.0;
  for (int base = 0; base < ni; base += 64) {
    const int q = base + lane;
    if (q < ni) {
      const int k = nbi[q];
      if (k != j) {
        Fs2Geo G;
        fs2_geometry(D, k, i, j, G);
        const double dkj2 = G.xkj[0] * G.xkj[0] + G.xkj[1] * G.xkj[1] + G.xkj[2] * G.xkj[2];
        if (dkj2 < K.rc2_2) {
          double sk[NS2][3];
          for (int s = 0; s < NS2; ++s)
            for (int c = 0; c < 3; ++c) sk[s][c] = D.site[((long)k * NS3 + s) * 3 + c];
          double dTdRi[3], dTdRj[3], Tl;
          const double T = fs2_T<false, true>(G, sk, nullptr, dTdRi, dTdRj, Tl);
          double fki, fpki, fkj, fpkj;
          fs2_damp(G.Rki, K, fki, fpki);
          fs2_damp(G.Rkj, K, fkj, fpkj);
          const double ff = fki * fkj;
          WT += ff * T;
          for (int c = 0; c < 3; ++c) {
            Fi[c] += -ff * S * dTdRi[c] + S * T * fpki * fkj * G.xki[c] / G.Rki;
            Fj[c] += -ff * S * dTdRj[c] + S * T * fki * fpkj * G.xkj[c] / G.Rkj;
          }
        }
      }
    }
  }
  WT = wave_sum(WT);
  for (int c = 0; c < 3; ++c) {
    Fi[c] = wave_sum(Fi[c]);
    Fj[c] = wave_sum(Fj[c]);
  }
  double* slot = D.pslot2 + p * PSLOT2;
  if (lane == 0) {
    slot[0] = S;
    slot[1] = Sl;
    for (int c = 0; c < 3; ++c) {
      slot[2 + 9 + c] = Fi[c];
      slot[14 + 9 + c] = Fj[c];
    }
  }
  if (lane < 3) {  // site forces of i and j: F = -WT dS/ds
    for (int c = 0; c < 3; ++c) {
      slot[2 + lane * 3 + c] = -WT * dSi[c];
      slot[14 + lane * 3 + c] = -WT * dSj[c];
    }
  }
}

// Spectator pass: one wavefront per spectator k over its items (k-major list).
// Output spec2[k] = [site forces of sites 0..4 (15), COM force (3), E, W].
__global__ __launch_bounds__(256) void k_fs2_spec(Dev D, Consts K) {
  const int k = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (k >= D.nm) return;
  const int s0 = D.off2[k], s1 = D.off2[k + 1];
  double sk[NS2][3];
  for (int s = 0; s < NS2; ++s)
    for (int c = 0; c < 3; ++c) sk[s][c] = D.site[((long)k * NS3 + s) * 3 + c];
  double acc[SPEC2];
  for (int a = 0; a < SPEC2; ++a) acc[a] = 0.0;
  for (int base = s0; base < s1; base += 64) {
    const int t = base + lane;
    if (t < s1) {
      const int4 it = D.items2[t];
      const int i = it.y, j = it.z;
      const double S = D.pslot2[(long)it.w * PSLOT2 + 0];
      const double Sl = D.pslot2[(long)it.w * PSLOT2 + 1];
      Fs2Geo G;
      fs2_geometry(D, k, i, j, G);
      double dTdsk[NS2][3], Tl;
      const double T = fs2_T<true, false>(G, sk, dTdsk, nullptr, nullptr, Tl);
      double fki, fpki, fkj, fpkj;
      fs2_damp(G.Rki, K, fki, fpki);
      fs2_damp(G.Rkj, K, fkj, fpkj);
      const double ff = fki * fkj;
      const double E = ff * S * T;
      for (int s = 0; s < NS2; ++s)
        for (int c = 0; c < 3; ++c) acc[s * 3 + c] -= ff * S * dTdsk[s][c];
      for (int c = 0; c < 3; ++c)
        acc[15 + c] -= S * T * (fpki * fkj * G.xki[c] / G.Rki + fki * fpkj * G.xkj[c] / G.Rkj);
      acc[18] += E;
      acc[19] += ff * (Sl * T + S * Tl) + S * T * (G.Rki * fpki * fkj + G.Rkj * fki * fpkj);
    }
  }
  for (int a = 0; a < SPEC2; ++a) acc[a] = wave_sum(acc[a]);
  if (lane < SPEC2) {
    double v = 0.0;
    for (int a = 0; a < SPEC2; ++a) v = a == lane ? acc[a] : v;
    D.spec2[(long)k * SPEC2 + lane] = v;
  }
}

__device__ __forceinline__ int find_sorted(const int* list, int n, int v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (list[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// FS2 gather: one wavefront per molecule; lanes 0..17 own (sites 0..4, COM).
// Order: spectator result, pairs (m, j) ascending j, pairs (i, m) ascending i.
__global__ __launch_bounds__(256) void k_fs2_gather(Dev D, int ld, double* fxs, double* fys,
                                                     double* fzs, double* fcom) {
  const int m = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (m >= D.nm) return;
  double acc = lane < 18 ? D.spec2[(long)m * SPEC2 + lane] : 0.0;
  // slot component feeding this lane (sites 0..2 and COM only)
  int sc = -1;
  if (lane < 9) sc = lane;
  else if (lane >= 15 && lane < 18) sc = 9 + (lane - 15);
  const int p0 = D.poff2[m], p1 = D.poff2[m + 1];
  for (int p = p0; p < p1; ++p)
    if (sc >= 0) acc += D.pslot2[(long)p * PSLOT2 + 2 + sc];
  const int* nb = D.nb2 + (long)m * D.cap_nb;
  const int n = D.nb2_cnt[m];
  for (int a = 0; a < n && nb[a] < m; ++a) {
    const int i = nb[a];
    const int* nbi = D.nb2 + (long)i * D.cap_nb;
    const int ni = D.nb2_cnt[i];
    const int u0 = find_sorted(nbi, ni, i + 1);
    const int pos = find_sorted(nbi, ni, m);
    const long p = (long)D.poff2[i] + (pos - u0);
    if (sc >= 0) acc += D.pslot2[p * PSLOT2 + 14 + sc];
  }
  const double conv = H2KCAL * 418.4 / BOHR2A;
  if (lane < 15) {
    const int s = lane / 3, c = lane - s * 3;
    double* f = c == 0 ? fxs : c == 1 ? fys : fzs;
    f[(long)ld * m + s] += acc * conv;
  } else if (lane < 18) {
    fcom[(long)(lane - 15) * D.nm + m] = acc * conv;
  }
}

// Per-box totals: {qfs2 (kcal/mol), qfs3 (kcal/mol), vir increment (10 J/mol)}.
__global__ __launch_bounds__(256) void k_energy(Dev D, int do2, int do3, double* out) {
  __shared__ double red[4];
  const int b = blockIdx.x;
  const int m0 = D.box_start[b], m1 = D.box_start[b + 1];
  double e2 = 0.0, w2 = 0.0, e3 = 0.0, w3 = 0.0;
  if (do2)
    for (int m = m0 + threadIdx.x; m < m1; m += 256) {
      e2 += D.spec2[(long)m * SPEC2 + 18];
      w2 += D.spec2[(long)m * SPEC2 + 19];
    }
  if (do3) {
    const long t0 = D.off3[m0], t1 = D.off3[m1];
    for (long t = t0 + threadIdx.x; t < t1; t += 256) {
      e3 += D.ew3[2 * t];
      w3 += D.ew3[2 * t + 1];
    }
  }
  e2 = block_sum256(e2, red);
  w2 = block_sum256(w2, red);
  e3 = block_sum256(e3, red);
  w3 = block_sum256(w3, red);
  if (threadIdx.x == 0) {
    out[3 * b + 0] = e2 * H2KCAL;
    out[3 * b + 1] = H2KCAL * e3;
    out[3 * b + 2] = w2 * H2KCAL * 418.4 + w3 * H2KCAL * 418.4;
  }
}

// explicit instantiations
template __global__ void k_fs3_eval<double>(Dev, Consts, const double4*, const double*);
template __global__ void k_fs3_eval<float>(Dev, Consts, const double4*, const float*);

}  // namespace ccpol3b
