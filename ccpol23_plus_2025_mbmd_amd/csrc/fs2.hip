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
// Two owner-computes passes, a few blocks per molecule each, so no force is
// scattered per item:
//   pair pass (blocks of i):      the pairs (i, j > i) and all their spectators
//                                 k; S_ij, the COM forces of i and j and the
//                                 exchange-site forces of both O,H,H triples;
//   spectator pass (blocks of k): all items with spectator k; the forces on k's
//                                 five sites and COM, and the item energies and
//                                 virials (summed per k in a fixed order).
// Both passes evaluate T_k, each with the derivatives it owns.  The acceptance
// test is the list kernels' arithmetic (no FMA contraction) on the minimum-image
// vectors the neighbour kernel stored with the FS2 lists, so both passes see
// exactly the reference's items.  Each block first tests its candidates and
// appends the accepted ones to an LDS queue in candidate order; the expensive T
// evaluation then runs over the queue with every lane busy (about 45 % of the
// candidates of a bulk-water box fail the derived-vector test).
#include "devcommon.h"

namespace ccpol3b {

__constant__ double c_q2[NS2] = {0.258504607653328389, 0.564050362569283648, 0.564050362569283648,
                                 -0.693302666395947953, -0.693302666395947953};
constexpr double ETA3 = 0.48056977605478400;
constexpr double BETA3 = 0.85836606579445000;
// The exchange-charge shift g of :177 is declared `real(8), parameter :: g3 =
// 0.400000000000000` without a kind suffix, so the literal is default REAL(4)
// (source/Makefile builds with `mpifort -O3`, no -fdefault-real-8): g3 holds
// the single-precision value of 0.4 widened to double, 0.4000000059604645.
constexpr double G3 = (double)0.4f;

struct Fs2Geo {
  double xij[3], xki[3], xkj[3];
  double Rij, Rki, Rkj;
};

__device__ __forceinline__ void fs2_damp(double R, const Consts& K, double& f, double& fp) {
  const double pom = exp(K.alp3 * (R - K.r0dmp2));
  const double pp = rcp_nr(1.0 + pom);
  f = pow_delta(pp, K);
  fp = -K.alp3 * K.delta33 * pom * pp * f;  // df/dR
}

// T_k and its derivatives.  sk: spectator site offsets [5][3].
// WANT_SITES: Fk[3 ks + c] += mult * dT/d(site ks)_c (the site derivatives are
// used at once, so they need no registers beyond the current site);
// dTdRi, dTdRj: explicit COM derivatives; Tl: dT/dlambda under uniform scaling
// of all vectors except the fixed shift g.  Inverse distances use rsqrt_nr
// (hardware estimate and two Newton steps) and one reciprocal of R_ij.
template <bool WANT_SITES, bool WANT_COM>
__device__ __forceinline__ double fs2_T(const Fs2Geo& G, const double* __restrict__ sk, double* Fk, double mult,
                                        double* dTdRi, double* dTdRj, double& Tl) {
  const double iRij = rcp_nr(G.Rij);
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
    const double ici = rsqrt_nr(dci[0] * dci[0] + dci[1] * dci[1] + dci[2] * dci[2]);
    const double icj = rsqrt_nr(dcj[0] * dcj[0] + dcj[1] * dcj[1] + dcj[2] * dcj[2]);
    const double iqi = rsqrt_nr(dqi[0] * dqi[0] + dqi[1] * dqi[1] + dqi[2] * dqi[2]);
    const double iqj = rsqrt_nr(dqj[0] * dqj[0] + dqj[1] * dqj[1] + dqj[2] * dqj[2]);
    const double qci3 = q * ici * ici * ici, qcj3 = q * icj * icj * icj;
    const double qqi3 = q * iqi * iqi * iqi, qqj3 = q * iqj * iqj * iqj;
    const double term0 = ici + icj - iqi - iqj;
    T += q * term0;
    const double pqi = dqi[0] * uij[0] + dqi[1] * uij[1] + dqi[2] * uij[2];
    const double pqj = dqj[0] * uij[0] + dqj[1] * uij[1] + dqj[2] * uij[2];
    Tl += -q * term0 + G3 * (pqi * qqi3 - pqj * qqj3);
    if constexpr (WANT_SITES)
      for (int c = 0; c < 3; ++c)
        Fk[ks * 3 + c] += mult * (-dci[c] * qci3 - dcj[c] * qcj3 + dqi[c] * qqi3 + dqj[c] * qqj3);
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

// The same T_k, site and COM derivatives with the charge-site distances in
// FP32 (the mixed precision mode).  The two inverse distances of a charge site
// to a molecule centre and to its exchange charge nearly cancel (the exchange
// charges sit g = 0.4 bohr from the centres), so each pair difference is formed
// without a subtraction: with c the vector from the centre, w the shift and
// q = c -/+ w, delta = |q|^2 - |c|^2 = g^2 -/+ 2 c.w and h = 1 / (|c| + |q|),
//   1/|c|   - 1/|q|   = delta h / (|c| |q|),
//   1/|c|^3 - 1/|q|^3 = delta h (|c|^2 + |c||q| + |q|^2) / (|c|^3 |q|^3).
// The displacements (FP64 sums) are rounded to FP32 once per site; the sums
// over sites, and everything outside T, stay FP64.
template <bool WANT_SITES, bool WANT_COM>
__device__ __forceinline__ double fs2_T_mixed(const Fs2Geo& G, const double* __restrict__ sk, double* Fk,
                                              double mult, double* dTdRi, double* dTdRj, double& Tl) {
  const double iRij = rcp_nr(G.Rij);
  const float u[3] = {float(G.xij[0] * iRij), float(G.xij[1] * iRij), float(G.xij[2] * iRij)};
  const float g = float(G3), g2 = g * g;  // G3 is a float value, so g is exact
  const float w[3] = {g * u[0], g * u[1], g * u[2]};
  const float gR = float(G3 * iRij);
  double T = 0.0;
  Tl = 0.0;
  if constexpr (WANT_COM)
    for (int c = 0; c < 3; ++c) dTdRi[c] = dTdRj[c] = 0.0;
#pragma unroll WANT_SITES ? NS2 : 1
  for (int ks = 0; ks < NS2; ++ks) {
    const float q = float(c_q2[ks]);
    float ci[3], cj[3];
    for (int c = 0; c < 3; ++c) {
      ci[c] = float(G.xki[c] + sk[ks * 3 + c]);
      cj[c] = float(G.xkj[c] + sk[ks * 3 + c]);
    }
    const float ci2 = ci[0] * ci[0] + ci[1] * ci[1] + ci[2] * ci[2];
    const float cj2 = cj[0] * cj[0] + cj[1] * cj[1] + cj[2] * cj[2];
    const float cui = ci[0] * u[0] + ci[1] * u[1] + ci[2] * u[2];
    const float cuj = cj[0] * u[0] + cj[1] * u[1] + cj[2] * u[2];
    const float di = g2 - 2.0f * g * cui;  // |q_i|^2 - |c_i|^2, q_i = c_i - w
    const float dj = g2 + 2.0f * g * cuj;  // |q_j|^2 - |c_j|^2, q_j = c_j + w
    const float qi2 = ci2 + di, qj2 = cj2 + dj;
    const float ici = __builtin_amdgcn_rsqf(ci2), icj = __builtin_amdgcn_rsqf(cj2);
    const float iqi = __builtin_amdgcn_rsqf(qi2), iqj = __builtin_amdgcn_rsqf(qj2);
    const float rci = ci2 * ici, rcj = cj2 * icj, rqi = qi2 * iqi, rqj = qj2 * iqj;
    const float hi = __builtin_amdgcn_rcpf(rci + rqi), hj = __builtin_amdgcn_rcpf(rcj + rqj);
    const float term0 = di * hi * ici * iqi + dj * hj * icj * iqj;  // 1/|c_i| - 1/|q_i| + 1/|c_j| - 1/|q_j|
    const float iqi3 = iqi * iqi * iqi, iqj3 = iqj * iqj * iqj;
    const float pqi = cui - g, pqj = cuj + g;  // q_i.u, q_j.u
    T += double(q * term0);
    Tl += double(-q * term0 + g * q * (pqi * iqi3 - pqj * iqj3));
    if constexpr (WANT_SITES || WANT_COM) {
      const float Bi = di * hi * (ci2 + rci * rqi + qi2) * (ici * ici * ici) * iqi3;  // 1/|c_i|^3 - 1/|q_i|^3
      const float Bj = dj * hj * (cj2 + rcj * rqj + qj2) * (icj * icj * icj) * iqj3;
      if constexpr (WANT_SITES)
        for (int c = 0; c < 3; ++c)
          Fk[ks * 3 + c] += mult * double(q * (-ci[c] * Bi - cj[c] * Bj + w[c] * (iqj3 - iqi3)));
      if constexpr (WANT_COM)
        for (int c = 0; c < 3; ++c) {
          const float Pqi = gR * (ci[c] - w[c] - pqi * u[c]);
          const float Pqj = gR * (cj[c] + w[c] - pqj * u[c]);
          const float x = q * (Pqj * iqj3 - Pqi * iqi3);
          dTdRi[c] += double(q * (ci[c] * Bi + w[c] * iqi3) + x);
          dTdRj[c] += double(q * (cj[c] * Bj - w[c] * iqj3) - x);
        }
    }
  }
  return T;
}

template <bool MIXED, bool WANT_SITES, bool WANT_COM>
__device__ __forceinline__ double fs2_T_in(const Fs2Geo& G, const double* __restrict__ sk, double* Fk, double mult,
                                           double* dTdRi, double* dTdRj, double& Tl) {
  if constexpr (MIXED) return fs2_T_mixed<WANT_SITES, WANT_COM>(G, sk, Fk, mult, dTdRi, dTdRj, Tl);
  else return fs2_T<WANT_SITES, WANT_COM>(G, sk, Fk, mult, dTdRi, dTdRj, Tl);
}

// Accepted candidates queued per block before they are evaluated.
constexpr int FS2_QCAP = 1024;
// The spectator pass takes FS2_SU candidates per thread at a time (their
// sub-list searches and loads in flight together) into a queue of FS2_SQCAP.
constexpr int FS2_SU = 4;
constexpr int FS2_SQCAP = 2048;

// Dynamic LDS of the spectator pass: 6 doubles and 4 ints per neighbour slot, the queue.
size_t fs2_lds_bytes(int cap_nb) {
  return (size_t)cap_nb * (6 * sizeof(double) + 4 * sizeof(int)) + FS2_SQCAP * sizeof(int) + 16;
}
// Dynamic LDS of the pair pass: 6 doubles and 1 int per neighbour slot; per pair
// slot S_ij and 8 fixed-point words (f f T, the COM force on j); the queue.
size_t fs2_pair_lds_bytes(int cap_nb) {
  return (size_t)cap_nb * (15 * sizeof(double) + sizeof(int)) + FS2_QCAP * sizeof(int) + 16;
}

// ---------------------------------------------------------------- pair pass
// D.fs2_pair_parts blocks per molecule i, each taking an equal share of i's
// upper pairs (i, j = nb2(i)[q]), q >= up2(i).  The candidates of pair q are the
// other entries k = nb2(i)[a], a != q; item (k; i, j) is accepted when the
// derived vector x_kj = x_ki + x_ij is inside the cutoff (:317-328).  Per pair
// the block keeps S_ij (:361-372), sum_k f f T (for the exchange-site forces
// F = -(sum_k f f T) dS/ds, :477-519) and the COM force on j in LDS; i's COM
// force stays in registers until the end.
template <bool MIXED>
__global__ __launch_bounds__(256) void k_fs2_pair_pass(Dev D, Consts K) {
  extern __shared__ double dsm[];
  __shared__ int wcnt[4];
  __shared__ unsigned long long sFsi[9][2];  // exchange-site forces on i, fixed point
  __shared__ double wred[4][3];
  const int parts = D.fs2_pair_parts;
  const int i = blockIdx.x / parts, part = blockIdx.x - i * parts;
  if (i >= D.nm) return;
  const int n = D.nb2_cnt[i], up = D.up2[i];
  const int share = (n - up + parts - 1) / parts;
  const int q0 = up + part * share, q1 = min(n, q0 + share);
  if (q0 >= q1) return;  // the whole block: no pair of i in this share
  const int npb = q1 - q0;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int cap = D.cap_nb;
  double* sx = dsm;  // x_ki = x_k - x_i (min image) of neighbour slot a
  double* sy = sx + cap;
  double* sz = sy + cap;
  double* sR = sz + cap;   // R_ki
  double* sF = sR + cap;   // f(R_ki)
  double* sFp = sF + cap;  // f'(R_ki)
  double* sS = sFp + cap;  // S_ij of pair slot ql = q - q0
  // per pair slot, summed over its accepted spectators k in fixed point (so
  // the order of the LDS atomics does not matter): f(R_ki) f(R_kj) T_k, and
  // the COM force on j
  unsigned long long* sWT = reinterpret_cast<unsigned long long*>(sS + cap);  // [cap][2]
  unsigned long long* sFj = sWT + 2 * cap;                                     // [cap][3][2]
  int* snb = reinterpret_cast<int*>(sFj + 6 * cap);
  int* queue = snb + cap;
  const int* nb = D.nb2 + (long)i * cap;
  const double* vx = D.nbv2 + (long)i * 3 * cap;
  const double* vy = vx + cap;
  const double* vz = vy + cap;
  const double* si = D.site + (long)i * NS3 * 3;
  for (int a = tid; a < n; a += 256) {
    snb[a] = nb[a];
    const double x = -vx[a], y = -vy[a], z = -vz[a];  // exact: minimg(-d) = -minimg(d)
    sx[a] = x;
    sy[a] = y;
    sz[a] = z;
    const double R = sqrt(x * x + y * y + z * z);
    sR[a] = R;
    fs2_damp(R, K, sF[a], sFp[a]);
  }
  if (tid < 9) sFsi[tid][0] = sFsi[tid][1] = 0ull;
  __syncthreads();
  // pair slots: S_ij and dS/dlambda (nine O,H,H site pairs, summed in order)
  for (int ql = tid; ql < npb; ql += 256) {
    const int q = q0 + ql;
    const double* sj = D.site + (long)snb[q] * NS3 * 3;
    const double xij[3] = {-sx[q], -sy[q], -sz[q]};
    double S = 0.0, Sl = 0.0;
    for (int is = 0; is < 3; ++is)
      for (int js = 0; js < 3; ++js) {
        double v[3];
        for (int c = 0; c < 3; ++c) v[c] = xij[c] + (si[is * 3 + c] - sj[js * 3 + c]);
        const double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        const double e = ETA3 * exp(-BETA3 * r);
        S += e;
        Sl += -BETA3 * r * e;
      }
    sS[ql] = S;
    for (int w = 0; w < 2; ++w) sWT[2 * ql + w] = 0ull;
    for (int w = 0; w < 6; ++w) sFj[6 * ql + w] = 0ull;
    const long p = (long)D.poff2[i] + (q - up);
    D.s2pair[2 * p] = S;
    D.s2pair[2 * p + 1] = Sl;
  }
  __syncthreads();
  // candidates (ql, a) in order; accepted ones are queued, then evaluated
  double Fi[3] = {0.0, 0.0, 0.0};
  const int ncand = npb * n;
  int qn = 0;
  for (int c0 = 0; c0 < ncand; c0 += 256) {
    const int cidx = c0 + tid;
    bool ok = false;
    int val = 0;
    if (cidx < ncand) {
      const int ql = cidx / n, a = cidx - ql * n;
      const int q = q0 + ql;
      if (a != q) {
        // x_kj = x_ki + x_ij with x_ij = -(slot q)
        ok = sq3_nc(sx[a] - sx[q], sy[a] - sy[q], sz[a] - sz[q]) < K.rc2_2;
        val = (ql << 16) | a;
      }
    }
    qn += block_append(ok, val, queue, qn, wcnt);
    if (qn <= FS2_QCAP - 256 && c0 + 256 < ncand) continue;  // block-uniform
    for (int e = tid; e < qn; e += 256) {
      const int ql = queue[e] >> 16, a = queue[e] & 0xffff;
      const int q = q0 + ql;
      Fs2Geo G;
      G.xij[0] = -sx[q];
      G.xij[1] = -sy[q];
      G.xij[2] = -sz[q];
      G.xki[0] = sx[a];
      G.xki[1] = sy[a];
      G.xki[2] = sz[a];
      G.xkj[0] = sx[a] - sx[q];
      G.xkj[1] = sy[a] - sy[q];
      G.xkj[2] = sz[a] - sz[q];
      G.Rij = sR[q];
      G.Rki = sR[a];
      const double r2kj = G.xkj[0] * G.xkj[0] + G.xkj[1] * G.xkj[1] + G.xkj[2] * G.xkj[2];
      const double iRkj = rsqrt_nr(r2kj);
      G.Rkj = r2kj * iRkj;
      double dTdRi[3], dTdRj[3], Tl;
      const double T =
          fs2_T_in<MIXED, false, true>(G, D.site + (long)snb[a] * NS3 * 3, nullptr, 0.0, dTdRi, dTdRj, Tl);
      const double fki = sF[a], fpki = sFp[a];
      double fkj, fpkj;
      fs2_damp(G.Rkj, K, fkj, fpkj);
      const double S = sS[ql];
      const double ff = fki * fkj, ffS = ff * S;
      const double ci = S * T * fpki * fkj * rcp_nr(G.Rki), cj = S * T * fki * fpkj * iRkj;
      for (int c = 0; c < 3; ++c) {
        Fi[c] += -ffS * dTdRi[c] + ci * G.xki[c];
        fx_atomic_add(sFj + 6 * ql + 2 * c, fx_checked(D, -ffS * dTdRj[c] + cj * G.xkj[c]));
      }
      fx_atomic_add(sWT + 2 * ql, fx_checked(D, ff * T));
    }
    qn = 0;
    __syncthreads();  // the queue is refilled
  }
  // exchange-site forces, F = (sum_k f f T) beta e_ab / r_ab v_ab on i's site a and
  // minus that on j's site b: six threads per pair, three summing over i's
  // sites for one site of j (and adding j's COM force component), three over
  // j's sites for one site of i
  for (int t = tid; t < npb * 6; t += 256) {
    const int ql = t / 6, role = t - 6 * ql;
    const int q = q0 + ql, j = snb[q];
    const double* sj = D.site + (long)j * NS3 * 3;
    const double xij[3] = {-sx[q], -sy[q], -sz[q]};
    const double WTb = fx_decode(sWT + 2 * ql) * BETA3;
    const int own = role < 3 ? role : role - 3;
    double f[3] = {0.0, 0.0, 0.0};
    for (int o = 0; o < 3; ++o) {
      const int is = role < 3 ? o : own, js = role < 3 ? own : o;
      double v[3];
      for (int c = 0; c < 3; ++c) v[c] = xij[c] + (si[is * 3 + c] - sj[js * 3 + c]);
      const double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      const double g = WTb * ETA3 * exp(-BETA3 * r) / r;
      for (int c = 0; c < 3; ++c) f[c] += g * v[c];
    }
    if (role < 3) {
      const long accj = (long)j * ACC_W;
      for (int c = 0; c < 3; ++c) fx_add(D, accj + own * 3 + c, -f[c]);
      const unsigned long long* fj = sFj + 6 * ql + 2 * own;  // already fixed point
      fx_atomic_add(D.acc + 2 * (accj + ACC_C2 + own), Fx128{fj[0], fj[1]});
    } else {
      for (int c = 0; c < 3; ++c) fx_atomic_add(sFsi[own * 3 + c], fx_checked(D, f[c]));
    }
  }
  for (int c = 0; c < 3; ++c) {
    const double v = wave_sum(Fi[c]);
    if (lane == 0) wred[w][c] = v;
  }
  __syncthreads();
  const long acci = (long)i * ACC_W;
  if (tid < 9) fx_atomic_add(D.acc + 2 * (acci + tid), Fx128{sFsi[tid][0], sFsi[tid][1]});
  else if (tid < 12) {
    const int c = tid - 9;
    fx_add(D, acci + ACC_C2 + c, (wred[0][c] + wred[1][c]) + (wred[2][c] + wred[3][c]));
  }
}

// ---------------------------------------------------------------- spectator pass
// D.fs2_parts blocks per molecule k, each taking an equal share of k's
// candidates (k; i = nb2(k)[a], j = nb2(i)[q]) with q >= up2(i), flattened in
// the reference's order (a, then q).  Each thread tests FS2_SU candidates at a
// time, 256 apart, and appends the accepted ones in a fixed order (thread by
// thread, each in its own candidate order); every thread then sums its queue
// entries in order, so the energy and virial sums are the same run to run.
template <bool MIXED>
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
  double* sx = dsm;  // x_ki = x_k - x_i (min image) of neighbour slot a
  double* sy = sx + cap;
  double* sz = sy + cap;
  double* sRk = sz + cap;
  double* sF = sRk + cap;
  double* sFp = sF + cap;
  int* nbl = reinterpret_cast<int*>(sFp + cap);
  int* off = nbl + cap;   // [cap+1] first candidate of sub-list a
  int* u0s = off + cap + 1;
  int* pofs = u0s + cap;  // pair index of candidate q of sub-list a: pofs[a] + q
  int* queue = pofs + cap;
  const double* vk = D.nbv2 + (long)k * 3 * cap;
  if (tid == 0) carry_s = 0;
  for (int a0 = 0; a0 < n; a0 += 256) {
    const int a = a0 + tid;
    int cnt = 0;
    if (a < n) {
      const int i = D.nb2[(long)k * cap + a];
      nbl[a] = i;
      const double x = vk[a], y = vk[cap + a], z = vk[2 * cap + a];
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
  const int f0 = min(total, part * share), f1 = min(total, f0 + share);
  const double* skp = D.site + (long)k * NS3 * 3;
  double Fk[NS2 * 3 + 3];
#pragma unroll
  for (int a = 0; a < NS2 * 3 + 3; ++a) Fk[a] = 0.0;
  double E = 0.0, W = 0.0;
  int nitems = 0, qn = 0;
  // sub-list search steps: powers of two from the largest below n
  int top = 0;
  if (n > 1) {
    top = 1;
    while (2 * top <= n - 1) top *= 2;
  }
  for (int base = f0; base < f1; base += 256 * FS2_SU) {
    // candidate f = base + 256 u + tid lies in sub-list lo[u]: the last a < n
    // with off[a] <= f (off is non-decreasing and off[0] = 0)
    int lo[FS2_SU];
#pragma unroll
    for (int u = 0; u < FS2_SU; ++u) lo[u] = 0;
    for (int sstep = top; sstep > 0; sstep >>= 1) {
#pragma unroll
      for (int u = 0; u < FS2_SU; ++u) {
        const int t = lo[u] + sstep;
        if (t < n && off[t] <= base + 256 * u + tid) lo[u] = t;
      }
    }
    bool ok[FS2_SU];
    int val[FS2_SU];
    int cnt = 0;
#pragma unroll
    for (int u = 0; u < FS2_SU; ++u) {
      const int f = base + 256 * u + tid;
      ok[u] = false;
      val[u] = 0;
      if (f < f1) {
        const int a = lo[u], i = nbl[a];
        const int qj = u0s[a] + (f - off[a]);
        const int jn = D.nb2[(long)i * cap + qj];
        const double* vi = D.nbv2 + (long)i * 3 * cap + qj;  // x_ij = x_i - x_j
        const bool in = sq3_nc(sx[a] + vi[0], sy[a] + vi[cap], sz[a] + vi[2 * cap]) < K.rc2_2;
        ok[u] = in & (jn != k);
        val[u] = (a << 16) | qj;
      }
      cnt += ok[u] ? 1 : 0;
    }
    // block exclusive scan of the per-thread counts; each thread appends its
    // accepted candidates in u order
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int v = __shfl_up(incl, d, 64);
      if (lane >= d) incl += v;
    }
    if (lane == 63) wtot[w] = incl;
    __syncthreads();
    int p = qn + incl - cnt;
    for (int u = 0; u < w; ++u) p += wtot[u];
#pragma unroll
    for (int u = 0; u < FS2_SU; ++u)
      if (ok[u]) queue[p++] = val[u];
    qn += (wtot[0] + wtot[1]) + (wtot[2] + wtot[3]);
    __syncthreads();
    if (qn <= FS2_SQCAP - 256 * FS2_SU && base + 256 * FS2_SU < f1) continue;  // block-uniform
    for (int e = tid; e < qn; e += 256) {
      const int a = queue[e] >> 16, qj = queue[e] & 0xffff;
      const int i = nbl[a];
      const double* vi = D.nbv2 + (long)i * 3 * cap + qj;
      Fs2Geo G;
      G.xki[0] = sx[a];
      G.xki[1] = sy[a];
      G.xki[2] = sz[a];
      G.xij[0] = vi[0];
      G.xij[1] = vi[cap];
      G.xij[2] = vi[2 * cap];
      G.xkj[0] = G.xki[0] + G.xij[0];
      G.xkj[1] = G.xki[1] + G.xij[1];
      G.xkj[2] = G.xki[2] + G.xij[2];
      ++nitems;
      const double r2ij = G.xij[0] * G.xij[0] + G.xij[1] * G.xij[1] + G.xij[2] * G.xij[2];
      G.Rij = r2ij * rsqrt_nr(r2ij);
      G.Rki = sRk[a];
      const double r2kj = G.xkj[0] * G.xkj[0] + G.xkj[1] * G.xkj[1] + G.xkj[2] * G.xkj[2];
      const double iRkj = rsqrt_nr(r2kj);
      G.Rkj = r2kj * iRkj;
      const long pr = (long)pofs[a] + qj;
      const double S = D.s2pair[2 * pr], Sl = D.s2pair[2 * pr + 1];
      const double fki = sF[a], fpki = sFp[a];
      double fkj, fpkj;
      fs2_damp(G.Rkj, K, fkj, fpkj);
      const double ff = fki * fkj, ffS = ff * S;
      double Tl;
      const double T = fs2_T_in<MIXED, true, false>(G, skp, Fk, -ffS, nullptr, nullptr, Tl);
      const double ci = S * T * fpki * fkj * rcp_nr(G.Rki), cj = S * T * fki * fpkj * iRkj;
      for (int c = 0; c < 3; ++c) Fk[NS2 * 3 + c] += -(ci * G.xki[c] + cj * G.xkj[c]);
      E += ff * S * T;
      W += ff * (Sl * T + S * Tl) + S * T * (G.Rki * fpki * fkj + G.Rkj * fki * fpkj);
    }
    qn = 0;
    __syncthreads();  // the queue is refilled
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
    const long acck = (long)k * ACC_W;
    const long e = (long)k * parts + part;
    if (tid < NS2 * 3) fx_add(D, acck + tid, v);
    else if (tid < NS2 * 3 + 3) fx_add(D, acck + ACC_C2 + (tid - NS2 * 3), v);
    else if (tid == NS2 * 3 + 3) D.ew2[2 * e] = v;
    else if (tid == NS2 * 3 + 4) D.ew2[2 * e + 1] = v;
    else if (v > 0.0) atomicAdd(D.totals + T_NI2, (int)v);
  }
}

template __global__ void k_fs2_pair_pass<false>(Dev, Consts);
template __global__ void k_fs2_pair_pass<true>(Dev, Consts);
template __global__ void k_fs2_spectator_pass<false>(Dev, Consts);
template __global__ void k_fs2_spectator_pass<true>(Dev, Consts);

}  // namespace ccpol3b
