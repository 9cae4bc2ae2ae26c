// FS2 (threebody_potential.f90:101-674) without a materialised item list.
//
// For an item (k; i<j) the reference energy is
//   E = f(R_ki) f(R_kj) S_ij T_k,  S_ij = eta sum_{a,b in O,H,H} exp(-beta r_ab),
//   T_k = sum_{ks=1..5} q_ks (1/|r_ks-R_i| + 1/|r_ks-R_j| - 1/|r_ks-R_i^+| - 1/|r_ks-R_j^+|)
// with R_i^+ = R_i + g u_ij, R_j^+ = R_j - g u_ij and f the damping of :339-349.
// The items are the reference's (:297-336): i an FS2 neighbour of k, j an upper
// FS2 neighbour of i other than k, accepted when the derived vector
// x_kj = x_ki + x_ij is inside the cutoff.
//
// Two owner-computes passes, one block per molecule each, so no force is
// scattered per item:
//   pair pass (block i):      the pairs (i, j) and all their spectators k;
//                             S_ij, the COM forces of i and j and the
//                             exchange-site forces of both O,H,H triples;
//   spectator pass (block k): all items with spectator k; the forces on k's
//                             five sites and COM, and the item energies and
//                             virials (summed per k in a fixed order).
// Both passes evaluate T_k, each with the derivatives it owns.  The acceptance
// test is the list kernels' arithmetic (no FMA contraction), so both passes see
// exactly the reference's items.
#include "devcommon.h"

namespace ccpol3b {

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
  f = pow_delta(pp, K);
  fp = -K.alp3 * K.delta33 * pom * pp * f;  // df/dR
}

// T_k and its derivatives.  sk: spectator site offsets [5][3].
// dTdsk: dT/d(site ks), dTdRi, dTdRj: explicit COM derivatives, Tl: dT/dlambda
// under uniform scaling of all vectors except the fixed shift g.  Inverse
// distances use the FP64 reciprocal square root (within 1 ulp) and one
// reciprocal of R_ij.
template <bool WANT_SITES, bool WANT_COM>
__device__ __forceinline__ double fs2_T(const Fs2Geo& G, const double* __restrict__ sk, double (*dTdsk)[3],
                                        double* dTdRi, double* dTdRj, double& Tl) {
  const double iRij = 1.0 / G.Rij;
  const double uij[3] = {G.xij[0] * iRij, G.xij[1] * iRij, G.xij[2] * iRij};
  const double gR = G3 * iRij;
  double T = 0.0;
  Tl = 0.0;
  if constexpr (WANT_COM)
    for (int c = 0; c < 3; ++c) dTdRi[c] = dTdRj[c] = 0.0;
  // the site loop stays rolled when only COM derivatives are wanted (registers)
#pragma unroll WANT_SITES ? NS2 : 1
  for (int ks = 0; ks < NS2; ++ks) {
    const double q = c_q2[ks];
    double dci[3], dcj[3], dqi[3], dqj[3];
    for (int c = 0; c < 3; ++c) {
      dci[c] = G.xki[c] + sk[ks * 3 + c];
      dcj[c] = G.xkj[c] + sk[ks * 3 + c];
      dqi[c] = dci[c] - G3 * uij[c];
      dqj[c] = dcj[c] + G3 * uij[c];
    }
    const double ici = rsqrt(dci[0] * dci[0] + dci[1] * dci[1] + dci[2] * dci[2]);
    const double icj = rsqrt(dcj[0] * dcj[0] + dcj[1] * dcj[1] + dcj[2] * dcj[2]);
    const double iqi = rsqrt(dqi[0] * dqi[0] + dqi[1] * dqi[1] + dqi[2] * dqi[2]);
    const double iqj = rsqrt(dqj[0] * dqj[0] + dqj[1] * dqj[1] + dqj[2] * dqj[2]);
    const double qci3 = q * ici * ici * ici, qcj3 = q * icj * icj * icj;
    const double qqi3 = q * iqi * iqi * iqi, qqj3 = q * iqj * iqj * iqj;
    const double term0 = ici + icj - iqi - iqj;
    T += q * term0;
    const double pqi = dqi[0] * uij[0] + dqi[1] * uij[1] + dqi[2] * uij[2];
    const double pqj = dqj[0] * uij[0] + dqj[1] * uij[1] + dqj[2] * uij[2];
    Tl += -q * term0 + G3 * (pqi * qqi3 - pqj * qqj3);
    if constexpr (WANT_SITES)
      for (int c = 0; c < 3; ++c)
        dTdsk[ks][c] = -dci[c] * qci3 - dcj[c] * qcj3 + dqi[c] * qqi3 + dqj[c] * qqj3;
    if constexpr (WANT_COM) {
      // projector P = I - u u^T applied to dqi, dqj
      for (int c = 0; c < 3; ++c) {
        const double Pqi = gR * (dqi[c] - pqi * uij[c]);
        const double Pqj = gR * (dqj[c] - pqj * uij[c]);
        dTdRi[c] += dci[c] * qci3 - (dqi[c] + Pqi) * qqi3 + Pqj * qqj3;
        dTdRj[c] += dcj[c] * qcj3 + Pqi * qqi3 - (dqj[c] + Pqj) * qqj3;
      }
    }
  }
  return T;
}

// Dynamic LDS of the spectator pass: 6 doubles and 4 ints per neighbour slot.
size_t fs2_lds_bytes(int cap_nb) { return (size_t)cap_nb * (6 * sizeof(double) + 4 * sizeof(int)) + 16; }

// ---------------------------------------------------------------- pair pass
__device__ void fs2_pair(const Dev& D, const Consts& K, long p, int lane);

// Wavefronts loop over the pairs p = (i, j = nb2(i)[q]), q >= up2(i); lanes run over
// the spectator candidates k = nb2(i)[a], a != q.  Lanes 0..8 also own the
// nine O,H,H site pairs of S_ij (:361-372) and of the exchange-site forces
// F = -(sum_k f f T) dS/ds (:477-519).
__global__ __launch_bounds__(256) void k_fs2_pair_pass(Dev D, Consts K) {
  const int lane = threadIdx.x & 63;
  const long np2 = D.totals[T_NP2];
  const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
  for (long p = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6; p < np2; p += nwaves)
    fs2_pair(D, K, p, lane);
}

// The pair p: everything it owns (see k_fs2_pair_pass).
__device__ void fs2_pair(const Dev& D, const Consts& K, long p, int lane) {
  const int i = owner_of(D.poff2, D.nm, p);
  const int n = D.nb2_cnt[i];
  const int jq = D.up2[i] + (int)(p - D.poff2[i]);
  const int* nb = D.nb2 + (long)i * D.cap_nb;
  const int j = nb[jq];
  const double L = D.boxL[D.mol_box[i]];
  const double xi = D.cx[i], yi = D.cy[i], zi = D.cz[i];
  const double xij[3] = {minimg_nc(xi - D.cx[j], L), minimg_nc(yi - D.cy[j], L), minimg_nc(zi - D.cz[j], L)};
  const double Rij = sqrt(xij[0] * xij[0] + xij[1] * xij[1] + xij[2] * xij[2]);
  // S_ij and its scaling derivative
  const double* si = D.site + (long)i * NS3 * 3;
  const double* sj = D.site + (long)j * NS3 * 3;
  const int is = lane / 3, js = lane - 3 * (lane / 3);
  double v9[3] = {0.0, 0.0, 0.0}, r9 = 1.0, e9 = 0.0, el9 = 0.0;
  if (lane < 9) {
    for (int c = 0; c < 3; ++c) v9[c] = xij[c] + (si[is * 3 + c] - sj[js * 3 + c]);
    r9 = sqrt(v9[0] * v9[0] + v9[1] * v9[1] + v9[2] * v9[2]);
    e9 = ETA3 * exp(-BETA3 * r9);
    el9 = -BETA3 * r9 * e9;
  }
  const double S = wave_sum(e9), Sl = wave_sum(el9);
  if (lane == 0) {
    D.s2pair[2 * p] = S;
    D.s2pair[2 * p + 1] = Sl;
  }
  // spectators
  double WT = 0.0, Fi[3] = {0.0, 0.0, 0.0}, Fj[3] = {0.0, 0.0, 0.0};
  for (int a = lane; a < n; a += 64) {
    if (a == jq) continue;
    const int k = nb[a];
    Fs2Geo G;
    G.xij[0] = xij[0];
    G.xij[1] = xij[1];
    G.xij[2] = xij[2];
    G.xki[0] = minimg_nc(D.cx[k] - xi, L);
    G.xki[1] = minimg_nc(D.cy[k] - yi, L);
    G.xki[2] = minimg_nc(D.cz[k] - zi, L);
    G.xkj[0] = G.xki[0] + xij[0];
    G.xkj[1] = G.xki[1] + xij[1];
    G.xkj[2] = G.xki[2] + xij[2];
    if (!(sq3_nc(G.xkj[0], G.xkj[1], G.xkj[2]) < K.rc2_2)) continue;
    G.Rij = Rij;
    G.Rki = sqrt(G.xki[0] * G.xki[0] + G.xki[1] * G.xki[1] + G.xki[2] * G.xki[2]);
    G.Rkj = sqrt(G.xkj[0] * G.xkj[0] + G.xkj[1] * G.xkj[1] + G.xkj[2] * G.xkj[2]);
    double dTdRi[3], dTdRj[3], Tl;
    const double T = fs2_T<false, true>(G, D.site + (long)k * NS3 * 3, nullptr, dTdRi, dTdRj, Tl);
    double fki, fpki, fkj, fpkj;
    fs2_damp(G.Rki, K, fki, fpki);
    fs2_damp(G.Rkj, K, fkj, fpkj);
    const double ff = fki * fkj, ffS = ff * S;
    const double ci = S * T * fpki * fkj / G.Rki, cj = S * T * fki * fpkj / G.Rkj;
    for (int c = 0; c < 3; ++c) {
      Fi[c] += -ffS * dTdRi[c] + ci * G.xki[c];
      Fj[c] += -ffS * dTdRj[c] + cj * G.xkj[c];
    }
    WT += ff * T;
  }
  WT = wave_sum(WT);
  for (int c = 0; c < 3; ++c) {
    Fi[c] = wave_sum(Fi[c]);
    Fj[c] = wave_sum(Fj[c]);
  }
  // exchange-site forces: lane (is, js) holds -WT dS/dv v-part; sum over js for
  // i's site is (lanes 3 is .. 3 is + 2) and over is for j's site js (lanes js, js+3, js+6)
  const double g1 = lane < 9 ? WT * BETA3 * e9 / r9 : 0.0;
  double* acci = D.acc + (long)i * ACC_W;
  double* accj = D.acc + (long)j * ACC_W;
  for (int c = 0; c < 3; ++c) {
    const double f = g1 * v9[c];
    const double fi = f + __shfl(f, lane + 1, 64) + __shfl(f, lane + 2, 64);
    const double fj = f + __shfl(f, lane + 3, 64) + __shfl(f, lane + 6, 64);
    if (lane < 9 && js == 0) unsafeAtomicAdd(acci + is * 3 + c, fi);
    if (lane < 3) unsafeAtomicAdd(accj + lane * 3 + c, -fj);
  }
  if (lane < 3) {
    unsafeAtomicAdd(acci + ACC_C2 + lane, lane == 0 ? Fi[0] : lane == 1 ? Fi[1] : Fi[2]);
    unsafeAtomicAdd(accj + ACC_C2 + lane, lane == 0 ? Fj[0] : lane == 1 ? Fj[1] : Fj[2]);
  }
}

// ---------------------------------------------------------------- spectator pass
// D.fs2_parts blocks per molecule k, each taking an equal share of k's items
// (k; i = nb2(k)[a], j = nb2(i)[q]) with q >= up2(i), flattened in the
// reference's order (a, then q).
__global__ __launch_bounds__(256) void k_fs2_spectator_pass(Dev D, Consts K) {
  extern __shared__ double dsm[];
  __shared__ int wtot[4];
  __shared__ int carry_s;
  const int parts = D.fs2_parts;
  const int k = blockIdx.x / parts, part = blockIdx.x - k * parts;
  if (k >= D.nm) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int n = D.nb2_cnt[k];
  const int cap = D.cap_nb;
  double* sx = dsm;  // x_ki = minimg(c_k - c_i) of neighbour slot a
  double* sy = sx + cap;
  double* sz = sy + cap;
  double* sRk = sz + cap;
  double* sF = sRk + cap;
  double* sFp = sF + cap;
  int* nbl = reinterpret_cast<int*>(sFp + cap);
  int* off = nbl + cap;   // [cap+1] first candidate of sub-list a
  int* u0s = off + cap + 1;
  int* pofs = u0s + cap;  // pair index of candidate q of sub-list a: pofs[a] + q
  const double L = D.boxL[D.mol_box[k]];
  const double xk = D.cx[k], yk = D.cy[k], zk = D.cz[k];
  if (tid == 0) carry_s = 0;
  for (int a0 = 0; a0 < n; a0 += 256) {
    const int a = a0 + tid;
    int cnt = 0;
    if (a < n) {
      const int i = D.nb2[(long)k * cap + a];
      nbl[a] = i;
      const double x = minimg_nc(xk - D.cx[i], L), y = minimg_nc(yk - D.cy[i], L), z = minimg_nc(zk - D.cz[i], L);
      sx[a] = x;
      sy[a] = y;
      sz[a] = z;
      const double R = sqrt(x * x + y * y + z * z);
      sRk[a] = R;
      fs2_damp(R, K, sF[a], sFp[a]);
      const int ui = D.up2[i];
      u0s[a] = ui;
      pofs[a] = D.poff2[i] - ui;
      cnt = D.nb2_cnt[i] - ui;
    }
    // block exclusive scan of the candidate counts
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int u = __shfl_up(incl, d, 64);
      if (lane >= d) incl += u;
    }
    if (lane == 63) wtot[w] = incl;
    __syncthreads();
    int wb = carry_s;
    for (int u = 0; u < w; ++u) wb += wtot[u];
    if (a < n) off[a] = wb + incl - cnt;
    __syncthreads();
    if (tid == 0) carry_s += (wtot[0] + wtot[1]) + (wtot[2] + wtot[3]);
    __syncthreads();
  }
  if (tid == 0) off[n] = carry_s;
  __syncthreads();
  const int total = off[n];
  const int share = (total + parts - 1) / parts;
  const int f0 = part * share, f1 = min(total, f0 + share);
  const double* skp = D.site + (long)k * NS3 * 3;
  double Fk[NS2 * 3 + 3];
#pragma unroll
  for (int a = 0; a < NS2 * 3 + 3; ++a) Fk[a] = 0.0;
  double E = 0.0, W = 0.0;
  int nitems = 0;
  for (int base = f0; base < f1; base += 256) {
    const int f = base + tid;
    if (f >= f1) continue;
    int lo = 0, hi = n;  // off[lo] <= f < off[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (off[mid] <= f) lo = mid; else hi = mid;
    }
    const int a = lo, i = nbl[a];
    const int qj = u0s[a] + (f - off[a]);
    const int j = D.nb2[(long)i * cap + qj];
    if (j == k) continue;
    Fs2Geo G;
    G.xki[0] = sx[a];
    G.xki[1] = sy[a];
    G.xki[2] = sz[a];
    G.xij[0] = minimg_nc(D.cx[i] - D.cx[j], L);
    G.xij[1] = minimg_nc(D.cy[i] - D.cy[j], L);
    G.xij[2] = minimg_nc(D.cz[i] - D.cz[j], L);
    G.xkj[0] = G.xki[0] + G.xij[0];
    G.xkj[1] = G.xki[1] + G.xij[1];
    G.xkj[2] = G.xki[2] + G.xij[2];
    if (!(sq3_nc(G.xkj[0], G.xkj[1], G.xkj[2]) < K.rc2_2)) continue;
    ++nitems;
    G.Rij = sqrt(G.xij[0] * G.xij[0] + G.xij[1] * G.xij[1] + G.xij[2] * G.xij[2]);
    G.Rki = sRk[a];
    G.Rkj = sqrt(G.xkj[0] * G.xkj[0] + G.xkj[1] * G.xkj[1] + G.xkj[2] * G.xkj[2]);
    const long pr = (long)pofs[a] + qj;
    const double S = D.s2pair[2 * pr], Sl = D.s2pair[2 * pr + 1];
    double dTdsk[NS2][3], Tl;
    const double T = fs2_T<true, false>(G, skp, dTdsk, nullptr, nullptr, Tl);
    const double fki = sF[a], fpki = sFp[a];
    double fkj, fpkj;
    fs2_damp(G.Rkj, K, fkj, fpkj);
    const double ff = fki * fkj, ffS = ff * S;
    for (int s2 = 0; s2 < NS2; ++s2)
      for (int c = 0; c < 3; ++c) Fk[s2 * 3 + c] += -ffS * dTdsk[s2][c];
    const double ci = S * T * fpki * fkj / G.Rki, cj = S * T * fki * fpkj / G.Rkj;
    for (int c = 0; c < 3; ++c) Fk[NS2 * 3 + c] += -(ci * G.xki[c] + cj * G.xkj[c]);
    E += ff * S * T;
    W += ff * (Sl * T + S * Tl) + S * T * (G.Rki * fpki * fkj + G.Rkj * fki * fpkj);
  }
  // fixed-order block sums of the 18 spectator force components, E, W and the
  // item count: wavefront butterflies, then one pass over the four wavefronts
  constexpr int NV = NS2 * 3 + 3 + 3;
  __shared__ double wred[4][NV];
  double vals[NV];
#pragma unroll
  for (int a = 0; a < NS2 * 3 + 3; ++a) vals[a] = Fk[a];
  vals[NS2 * 3 + 3] = E;
  vals[NS2 * 3 + 4] = W;
  vals[NS2 * 3 + 5] = (double)nitems;
#pragma unroll
  for (int a = 0; a < NV; ++a) {
    const double v = wave_sum(vals[a]);
    if (lane == 0) wred[w][a] = v;
  }
  __syncthreads();
  if (tid < NV) {
    const double v = (wred[0][tid] + wred[1][tid]) + (wred[2][tid] + wred[3][tid]);
    double* acck = D.acc + (long)k * ACC_W;
    const long e = (long)k * parts + part;
    if (tid < NS2 * 3) unsafeAtomicAdd(acck + tid, v);
    else if (tid < NS2 * 3 + 3) unsafeAtomicAdd(acck + ACC_C2 + (tid - NS2 * 3), v);
    else if (tid == NS2 * 3 + 3) D.ew2[2 * e] = v;
    else if (tid == NS2 * 3 + 4) D.ew2[2 * e + 1] = v;
    else if (v > 0.0) atomicAdd(D.totals + T_NI2, (int)v);
  }
}

}  // namespace ccpol3b
