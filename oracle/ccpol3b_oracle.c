/*
 * ccpol3b_oracle.c -- CPU FP64 oracle for the CCpol23+(2025) three-body term.
 *
 * TEST INFRASTRUCTURE ONLY (see ccpol3b_oracle.h).  A plain-C restatement of
 * source/threebody_potential.f90 (threebody_term :17-99, FS2_term :101-674,
 * FS3_term :676-2062, fill :2064-2131, add_pair_virial_trace :2133-2142).
 *
 * Faithfulness rules (they matter for the bit-exact triplet lists):
 *   - built with -ffp-contract=off and without -ffast-math, like the
 *     reference's x86-64 `mpifort -O3` build, which has no FMA;
 *   - Fortran NINT is round-half-away-from-zero -> C round();
 *   - sums are evaluated left to right as written in the reference;
 *   - constant exponents with real kind (x**2.d0, bohr2a**(-2.d0)) go through
 *     libm pow, the rest through multiplication.
 * Accumulation order follows the reference loop nests when nthreads == 1.
 *
 * With orc_opts.optimised the same routine is the optimised CPU baseline the
 * benchmark reports beside the faithful one: neighbour-list list builders
 * (fs3_list_nb, fs2_list_nb) and the factorised FS3 region
 * (fs3_region_canonical); tests/test_oracle_canonical.py checks it against the
 * restatement.
 */
#include "ccpol3b_oracle.h"

#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* common_arrays.f90:34-35 */
static const double bohr2a = 0.529177249;
static const double h2kcal = 627.510;

/* FS3 site types, threebody_potential.f90:2076 */
static const int ntype3[ORC_NS3] = {1, 2, 2, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 6, 6, 6, 6};

/* FS2 constants, threebody_potential.f90:175-177, 197-201 */
static const double eta3 = 0.48056977605478400;
static const double beta3 = 0.85836606579445000;
/* g3 of :177 is written without a kind suffix (`0.400000000000000`), so it is
 * a default REAL(4) literal widened to real(8): 0.4f = 0.4000000059604645
 * (source/Makefile: mpifort -O3, no -fdefault-real-8). */
static const double g3 = (double)0.4f;
static const double charge3b[5] = {0.258504607653328389, 0.564050362569283648,
                                   0.564050362569283648, -0.693302666395947953,
                                   -0.693302666395947953};

static double wall_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void orc_default_opts(orc_opts *o) {
  memset(o, 0, sizeof(*o));
  o->nthreads = 1;
  o->fs3_first = 0;
  o->fs3_last = -1;
  o->fs2_first = 0;
  o->fs2_last = -1;
}

double orc_fs3_rc2_bohr2(void) {
  /* rc2_3b = 6.3d0**(2.d0)*bohr2a**(-2.d0)  (:770) */
  return pow(6.3, 2.0) * pow(bohr2a, -2.0);
}

double orc_fs2_rc2_bohr2(double rcut3) {
  /* rcutsd3 = rcut3**2 (initialization.f90:158); rc2_3b = rcutsd3*bohr2a**(-2.d0) (:190) */
  double rcutsd3 = rcut3 * rcut3;
  return rcutsd3 * pow(bohr2a, -2.0);
}

/* ------------------------------------------------------------------------ */
/* fill (:2064-2131): number every member of one orbit of
 * {(type(iat),ip),(type(jat),jp),(type(kat),kp)} with nr.  Layout
 * ntab(kp,jp,ip,kat,jat,iat), Fortran column-major, 1-based sites. */
static inline int tab_index(int kp, int jp, int ip, int kat, int jat, int iat) {
  /* idx-1 of :924-929 */
  return kp + 2 * jp + 4 * ip + 8 * (kat - 1) + 136 * (jat - 1) + 2312 * (iat - 1);
}

static void type_block(int site, int *first, int *last) {
  int t = ntype3[site - 1], s;
  *first = -1;
  *last = -1;
  for (s = 1; s <= ORC_NS3; ++s) {
    if (ntype3[s - 1] == t) {
      if (*first < 0) *first = s;
      *last = s;
    }
  }
}

static void fill_orbit(int *ntab, int iat, int jat, int kat, int ip, int jp, int kp, int nr) {
  int ifr, ila, jfr, jla, kfr, kla, ii, jj, kk;
  type_block(iat, &ifr, &ila);
  type_block(jat, &jfr, &jla);
  type_block(kat, &kfr, &kla);
  for (ii = ifr; ii <= ila; ++ii)
    for (jj = jfr; jj <= jla; ++jj)
      for (kk = kfr; kk <= kla; ++kk) {
        /* the six joint permutations of (site, power), :2121-2126 */
        ntab[tab_index(kp, jp, ip, kk, jj, ii)] = nr;
        ntab[tab_index(jp, kp, ip, jj, kk, ii)] = nr;
        ntab[tab_index(ip, kp, jp, ii, kk, jj)] = nr;
        ntab[tab_index(kp, ip, jp, kk, ii, jj)] = nr;
        ntab[tab_index(ip, jp, kp, ii, jj, kk)] = nr;
        ntab[tab_index(jp, ip, kp, jj, ii, kk)] = nr;
      }
}

int orc_build_ind_lin(int *ind_lin_2) {
  /* :909-946 -- first-encounter numbering in loop order iat,jat,kat,ip,jp,kp */
  int iat, jat, kat, ip, jp, kp, nnew = 0;
  memset(ind_lin_2, 0, sizeof(int) * ORC_NZERO);
  for (iat = 1; iat <= ORC_NS3; ++iat)
    for (jat = 1; jat <= ORC_NS3; ++jat)
      for (kat = 1; kat <= ORC_NS3; ++kat)
        for (ip = 0; ip <= 1; ++ip)
          for (jp = 0; jp <= 1; ++jp)
            for (kp = 0; kp <= 1; ++kp) {
              int idx = tab_index(kp, jp, ip, kat, jat, iat);
              if (ind_lin_2[idx] == 0) {
                ++nnew;
                fill_orbit(ind_lin_2, iat, jat, kat, ip, jp, kp, nnew);
              }
            }
  return nnew;
}

/* ------------------------------------------------------------------------ */
/* minimum image as written in the reference: d - boxi*nint(d/boxi) */
static inline double nint_d(double v) { return round(v); }

typedef struct {
  int nm, ld;
  const double *xx, *yy, *zz;     /* bohr, [nm] */
  const double *xxs, *yys, *zzs;  /* bohr, [ld*nm] */
  double boxi;                    /* bohr; <= 0 -> open boundary */
} geom_t;

static void to_bohr(int nm, int ld, int nsite_used,
                    const double *x, const double *y, const double *z,
                    const double *xs, const double *ys, const double *zs, double box,
                    double *xx, double *yy, double *zz,
                    double *xxs, double *yys, double *zzs, double *boxi) {
  int i, is;
  for (i = 0; i < nm; ++i) {
    xx[i] = x[i] / bohr2a;
    yy[i] = y[i] / bohr2a;
    zz[i] = z[i] / bohr2a;
    for (is = 0; is < nsite_used; ++is) {
      xxs[is + (long)ld * i] = xs[is + (long)ld * i] / bohr2a;
      yys[is + (long)ld * i] = ys[is + (long)ld * i] / bohr2a;
      zzs[is + (long)ld * i] = zs[is + (long)ld * i] / bohr2a;
    }
  }
  *boxi = box > 0.0 ? box / bohr2a : 0.0;
}

/* FS3 acceptance predicate and COM vectors (:862-892). */
static inline int fs3_accept(const geom_t *G, int i, int j, int k, double rc2,
                             double v[9], double d2[3]) {
  const double L = G->boxi;
  double dxij = G->xx[i] - G->xx[j], dyij = G->yy[i] - G->yy[j], dzij = G->zz[i] - G->zz[j];
  double dxik = G->xx[i] - G->xx[k], dyik = G->yy[i] - G->yy[k], dzik = G->zz[i] - G->zz[k];
  double dxjk = G->xx[j] - G->xx[k], dyjk = G->yy[j] - G->yy[k], dzjk = G->zz[j] - G->zz[k];
  double xij, yij, zij, xik, yik, zik, xjk, yjk, zjk;
  if (L > 0.0) {
    double nxij = nint_d(dxij / L), nyij = nint_d(dyij / L), nzij = nint_d(dzij / L);
    double nxik = nint_d(dxik / L), nyik = nint_d(dyik / L), nzik = nint_d(dzik / L);
    xij = dxij - L * nxij; yij = dyij - L * nyij; zij = dzij - L * nzij;
    xik = dxik - L * nxik; yik = dyik - L * nyik; zik = dzik - L * nzik;
    /* the derived third vector, evaluated left to right (:882-884) */
    xjk = dxjk + L * nxij - L * nxik;
    yjk = dyjk + L * nyij - L * nyik;
    zjk = dzjk + L * nzij - L * nzik;
  } else {
    xij = dxij; yij = dyij; zij = dzij;
    xik = dxik; yik = dyik; zik = dzik;
    xjk = dxjk; yjk = dyjk; zjk = dzjk;
  }
  d2[0] = xij * xij + yij * yij + zij * zij;
  d2[1] = xik * xik + yik * yik + zik * zik;
  d2[2] = xjk * xjk + yjk * yjk + zjk * zjk;
  v[0] = xij; v[1] = yij; v[2] = zij;
  v[3] = xik; v[4] = yik; v[5] = zik;
  v[6] = xjk; v[7] = yjk; v[8] = zjk;
  return (d2[0] < rc2) && (d2[1] < rc2) && (d2[2] < rc2);
}

/* FS2 acceptance predicate and COM vectors (:305-336). */
static inline int fs2_accept(const geom_t *G, int k, int i, int j, double rc2,
                             double v[9], double d2[3]) {
  const double L = G->boxi;
  double dxij = G->xx[i] - G->xx[j], dyij = G->yy[i] - G->yy[j], dzij = G->zz[i] - G->zz[j];
  double dxki = G->xx[k] - G->xx[i], dyki = G->yy[k] - G->yy[i], dzki = G->zz[k] - G->zz[i];
  double xij, yij, zij, xki, yki, zki, xkj, ykj, zkj;
  if (L > 0.0) {
    xij = dxij - L * nint_d(dxij / L);
    yij = dyij - L * nint_d(dyij / L);
    zij = dzij - L * nint_d(dzij / L);
    xki = dxki - L * nint_d(dxki / L);
    yki = dyki - L * nint_d(dyki / L);
    zki = dzki - L * nint_d(dzki / L);
  } else {
    xij = dxij; yij = dyij; zij = dzij;
    xki = dxki; yki = dyki; zki = dzki;
  }
  /* Keep all three COM vectors in the same periodic image (:325-328) */
  xkj = xki + xij;
  ykj = yki + yij;
  zkj = zki + zij;
  d2[0] = xij * xij + yij * yij + zij * zij;
  d2[1] = xki * xki + yki * yki + zki * zki;
  d2[2] = xkj * xkj + ykj * ykj + zkj * zkj;
  v[0] = xij; v[1] = yij; v[2] = zij;
  v[3] = xki; v[4] = yki; v[5] = zki;
  v[6] = xkj; v[7] = ykj; v[8] = zkj;
  return (d2[0] < rc2) && (d2[1] < rc2) && (d2[2] < rc2);
}

static long fs3_list_bohr(const geom_t *G, double rc2, int *out, long cap) {
  long n = 0;
  int i, j, k;
  double v[9], d2[3];
  for (i = 0; i <= G->nm - 3; ++i)
    for (j = i + 1; j <= G->nm - 2; ++j)
      for (k = j + 1; k <= G->nm - 1; ++k)
        if (fs3_accept(G, i, j, k, rc2, v, d2)) {
          if (out && n < cap) {
            out[3 * n] = i;
            out[3 * n + 1] = j;
            out[3 * n + 2] = k;
          }
          ++n;
        }
  return n;
}

static long fs2_list_bohr(const geom_t *G, double rc2, int *out, long cap) {
  long n = 0;
  int i, j, k;
  double v[9], d2[3];
  for (k = 0; k < G->nm; ++k)
    for (i = 0; i < G->nm - 1; ++i) {
      if (k == i) continue;
      for (j = i + 1; j < G->nm; ++j) {
        if (j == k) continue;
        if (fs2_accept(G, k, i, j, rc2, v, d2)) {
          if (out && n < cap) {
            out[3 * n] = k;
            out[3 * n + 1] = i;
            out[3 * n + 2] = j;
          }
          ++n;
        }
      }
    }
  return n;
}

/* The optimised baseline's list builders: the same predicates and loop order
 * as fs3_list_bohr / fs2_list_bohr, enumerated from neighbour lists.  A pair
 * fails the full predicate when it fails its own distance test, computed with
 * the same arithmetic, so only candidates from the lists can pass.  nb[m] holds
 * the molecules j != m with |minimg(x_m - x_j)|^2 < rc2 in ascending order
 * (minimg(-d) = -minimg(d) exactly, so the relation is symmetric). */
static int **neighbour_lists(const geom_t *G, double rc2, int **count_out) {
  const double L = G->boxi;
  int m, j;
  int **nb = (int **)calloc(G->nm ? G->nm : 1, sizeof(int *));
  int *cnt = (int *)calloc(G->nm ? G->nm : 1, sizeof(int));
  for (m = 0; m < G->nm; ++m) {
    int cap = 16;
    nb[m] = (int *)malloc(sizeof(int) * cap);
    for (j = 0; j < G->nm; ++j) {
      double dx = G->xx[m] - G->xx[j], dy = G->yy[m] - G->yy[j], dz = G->zz[m] - G->zz[j];
      if (j == m) continue;
      if (L > 0.0) {
        dx = dx - L * nint_d(dx / L);
        dy = dy - L * nint_d(dy / L);
        dz = dz - L * nint_d(dz / L);
      }
      if (dx * dx + dy * dy + dz * dz < rc2) {
        if (cnt[m] == cap) nb[m] = (int *)realloc(nb[m], sizeof(int) * (cap *= 2));
        nb[m][cnt[m]++] = j;
      }
    }
  }
  *count_out = cnt;
  return nb;
}

static void free_neighbour_lists(int **nb, int *cnt, int nm) {
  int m;
  for (m = 0; m < nm; ++m) free(nb[m]);
  free(nb);
  free(cnt);
}

static long fs3_list_nb(const geom_t *G, double rc2, int *out, long cap) {
  long n = 0;
  int i, a, b, *cnt;
  double v[9], d2[3];
  int **nb = neighbour_lists(G, rc2, &cnt);
  for (i = 0; i < G->nm; ++i)
    for (a = 0; a < cnt[i]; ++a) {
      const int j = nb[i][a];
      if (j <= i) continue;
      for (b = a + 1; b < cnt[i]; ++b) {
        const int k = nb[i][b];
        if (fs3_accept(G, i, j, k, rc2, v, d2)) {
          if (out && n < cap) {
            out[3 * n] = i;
            out[3 * n + 1] = j;
            out[3 * n + 2] = k;
          }
          ++n;
        }
      }
    }
  free_neighbour_lists(nb, cnt, G->nm);
  return n;
}

static long fs2_list_nb(const geom_t *G, double rc2, int *out, long cap) {
  long n = 0;
  int k, a, b, *cnt;
  double v[9], d2[3];
  int **nb = neighbour_lists(G, rc2, &cnt);
  for (k = 0; k < G->nm; ++k)
    for (a = 0; a < cnt[k]; ++a) {
      const int i = nb[k][a];
      for (b = 0; b < cnt[i]; ++b) {
        const int j = nb[i][b];
        if (j <= i || j == k) continue;
        if (fs2_accept(G, k, i, j, rc2, v, d2)) {
          if (out && n < cap) {
            out[3 * n] = k;
            out[3 * n + 1] = i;
            out[3 * n + 2] = j;
          }
          ++n;
        }
      }
    }
  free_neighbour_lists(nb, cnt, G->nm);
  return n;
}

typedef struct {
  double *x, *y, *z, *xs, *ys, *zs;
  double boxi;
} bohr_buf;

static int make_geom(int nm, int ld, int nsite_used, const double *x, const double *y,
                     const double *z, const double *xs, const double *ys, const double *zs,
                     double box, geom_t *G, bohr_buf *B) {
  size_t nsit = (size_t)ld * nm;
  B->x = (double *)calloc(nm ? nm : 1, sizeof(double));
  B->y = (double *)calloc(nm ? nm : 1, sizeof(double));
  B->z = (double *)calloc(nm ? nm : 1, sizeof(double));
  B->xs = (double *)calloc(nsit ? nsit : 1, sizeof(double));
  B->ys = (double *)calloc(nsit ? nsit : 1, sizeof(double));
  B->zs = (double *)calloc(nsit ? nsit : 1, sizeof(double));
  if (!B->x || !B->y || !B->z || !B->xs || !B->ys || !B->zs) return -1;
  to_bohr(nm, ld, nsite_used, x, y, z, xs ? xs : B->xs, ys ? ys : B->ys, zs ? zs : B->zs,
          box, B->x, B->y, B->z, B->xs, B->ys, B->zs, &B->boxi);
  G->nm = nm;
  G->ld = ld;
  G->xx = B->x; G->yy = B->y; G->zz = B->z;
  G->xxs = B->xs; G->yys = B->ys; G->zzs = B->zs;
  G->boxi = B->boxi;
  return 0;
}

static void free_bohr(bohr_buf *B) {
  free(B->x); free(B->y); free(B->z);
  free(B->xs); free(B->ys); free(B->zs);
}

long orc_fs3_list(const orc_params *P, int nm, const double *x, const double *y,
                  const double *z, double box, int *out, long cap) {
  geom_t G;
  bohr_buf B;
  long n;
  (void)P;
  if (make_geom(nm, 1, 0, x, y, z, NULL, NULL, NULL, box, &G, &B)) return -1;
  n = fs3_list_bohr(&G, orc_fs3_rc2_bohr2(), out, cap);
  free_bohr(&B);
  return n;
}

long orc_fs2_list(const orc_params *P, int nm, const double *x, const double *y,
                  const double *z, double box, int *out, long cap) {
  geom_t G;
  bohr_buf B;
  long n;
  if (make_geom(nm, 1, 0, x, y, z, NULL, NULL, NULL, box, &G, &B)) return -1;
  n = fs2_list_bohr(&G, orc_fs2_rc2_bohr2(P->rcut3), out, cap);
  free_bohr(&B);
  return n;
}

/* ------------------------------------------------------------------------ */
/* Accumulators of one worker (the reference's lfxs/lfxfs*, lqqfs*, lqvirial) */
typedef struct {
  double *fx, *fy, *fz;     /* [ld*nm] site forces, hartree/bohr */
  double *cx, *cy, *cz;     /* [nm] COM forces */
  double e, e0, w;          /* energy, undamped FS2 energy (qqfs20), virial */
} acc_t;

static int acc_alloc(acc_t *a, int nm, int ld) {
  size_t n = (size_t)ld * nm;
  a->fx = (double *)calloc(n ? n : 1, sizeof(double));
  a->fy = (double *)calloc(n ? n : 1, sizeof(double));
  a->fz = (double *)calloc(n ? n : 1, sizeof(double));
  a->cx = (double *)calloc(nm ? nm : 1, sizeof(double));
  a->cy = (double *)calloc(nm ? nm : 1, sizeof(double));
  a->cz = (double *)calloc(nm ? nm : 1, sizeof(double));
  a->e = a->e0 = a->w = 0.0;
  return (a->fx && a->fy && a->fz && a->cx && a->cy && a->cz) ? 0 : -1;
}

static void acc_free(acc_t *a) {
  free(a->fx); free(a->fy); free(a->fz);
  free(a->cx); free(a->cy); free(a->cz);
}

/* add_pair_virial_trace (:2133-2142) */
static inline void add_pair_virial_trace(double *q, double coeffr, double xr, double yr,
                                         double zr) {
  *q = *q - coeffr * (xr * xr + yr * yr + zr * zr);
}

/* One region sweep of FS3 (the 17^3 x 2^3 loops at :949-1180, :1183-1415,
 * :1418-1943).  The reference writes four copies that differ only in the
 * parameter tables and in the weights:
 *   energy/site weight wE (1, smfun or 1-smfun) and the COM radial factor
 *   wC (dfdmp, dfdmp*smfun+fdmp*cf or dfdmp*smfun2-fdmp*cf).               */
static void fs3_region(const geom_t *G, int i, int j, int k, const double v[9],
                       const double R[3], const double (*bS)[17], const double (*gS)[17],
                       const double (*r0S)[17], const double *cs, const int *tbl,
                       double fdmp, double wE, double wC, acc_t *A) {
  const int ld = G->ld;
  const double xij = v[0], yij = v[1], zij = v[2];
  const double xik = v[3], yik = v[4], zik = v[5];
  const double xjk = v[6], yjk = v[7], zjk = v[8];
  const double Rij = R[0], Rik = R[1], Rjk = R[2];
  const double *xxs = G->xxs, *yys = G->yys, *zzs = G->zzs;
  int is, js, ks, ip, jp, kp, itab = 0;
  for (is = 0; is < ORC_NS3; ++is)
    for (js = 0; js < ORC_NS3; ++js)
      for (ks = 0; ks < ORC_NS3; ++ks) {
        const long si = is + (long)ld * i, sj = js + (long)ld * j;
        const long sik = is + (long)ld * i, sk = ks + (long)ld * k;
        const long sjk = js + (long)ld * j;
        double bsij = bS[is][js], gsij = gS[is][js], r0sij = r0S[is][js];
        double xsij = xij + (xxs[si] - xxs[sj]);
        double ysij = yij + (yys[si] - yys[sj]);
        double zsij = zij + (zzs[si] - zzs[sj]);
        double dsij2 = xsij * xsij + ysij * ysij + zsij * zsij;
        double dsij = sqrt(dsij2);
        double dampsij = 1.0 / (1.0 + exp(-gsij * (dsij - r0sij)));

        double bsik = bS[is][ks], gsik = gS[is][ks], r0sik = r0S[is][ks];
        double xsik = xik + (xxs[sik] - xxs[sk]);
        double ysik = yik + (yys[sik] - yys[sk]);
        double zsik = zik + (zzs[sik] - zzs[sk]);
        double dsik2 = xsik * xsik + ysik * ysik + zsik * zsik;
        double dsik = sqrt(dsik2);
        double dampsik = 1.0 / (1.0 + exp(-gsik * (dsik - r0sik)));

        double bsjk = bS[js][ks], gsjk = gS[js][ks], r0sjk = r0S[js][ks];
        double xsjk = xjk + (xxs[sjk] - xxs[sk]);
        double ysjk = yjk + (yys[sjk] - yys[sk]);
        double zsjk = zjk + (zzs[sjk] - zzs[sk]);
        double dsjk2 = xsjk * xsjk + ysjk * ysjk + zsjk * zsjk;
        double dsjk = sqrt(dsjk2);
        double dampsjk = 1.0 / (1.0 + exp(-gsjk * (dsjk - r0sjk)));

        double eksij = exp(-bsij * dsij) * dampsij;
        double eksik = exp(-bsik * dsik) * dampsik;
        double eksjk = exp(-bsjk * dsjk) * dampsjk;
        double eks = eksij * eksik * eksjk;

        double der_fij = -bsij + gsij * exp(-gsij * (dsij - r0sij)) * dampsij;
        double der_fik = -bsik + gsik * exp(-gsik * (dsik - r0sik)) * dampsik;
        double der_fjk = -bsjk + gsjk * exp(-gsjk * (dsjk - r0sjk)) * dampsjk;

        for (ip = 0; ip <= 1; ++ip) {
          double term1 = ip ? eks * dsjk : eks;
          double fsjk = der_fjk / dsjk + (double)ip / (dsjk * dsjk);
          for (jp = 0; jp <= 1; ++jp) {
            double term2 = jp ? term1 * dsik : term1;
            double fsik = der_fik / dsik + (double)jp / (dsik * dsik);
            for (kp = 0; kp <= 1; ++kp) {
              int iind = tbl[itab++] - 1;
              double term3 = kp ? term2 * dsij : term2;
              double fsij = der_fij / dsij + (double)kp / (dsij * dsij);
              double c = cs[iind];
              double coeff;

              A->e = A->e + wE * (c * term3 * fdmp);

              coeff = -c * term3 * fdmp * wE * fsij;
              add_pair_virial_trace(&A->w, coeff, xsij, ysij, zsij);
              coeff = -c * term3 * fdmp * wE * fsik;
              add_pair_virial_trace(&A->w, coeff, xsik, ysik, zsik);
              coeff = -c * term3 * fdmp * wE * fsjk;
              add_pair_virial_trace(&A->w, coeff, xsjk, ysjk, zsjk);
              coeff = -c * term3 * wC / Rij;
              add_pair_virial_trace(&A->w, coeff, xij, yij, zij);
              coeff = -c * term3 * wC / Rik;
              add_pair_virial_trace(&A->w, coeff, xik, yik, zik);
              coeff = -c * term3 * wC / Rjk;
              add_pair_virial_trace(&A->w, coeff, xjk, yjk, zjk);

              /* molecule i */
              A->fx[si] -= c * term3 * (fsij * xsij + fsik * xsik) * fdmp * wE;
              A->fy[si] -= c * term3 * (fsij * ysij + fsik * ysik) * fdmp * wE;
              A->fz[si] -= c * term3 * (fsij * zsij + fsik * zsik) * fdmp * wE;
              A->cx[i] -= c * term3 * wC * (xij / Rij + xik / Rik);
              A->cy[i] -= c * term3 * wC * (yij / Rij + yik / Rik);
              A->cz[i] -= c * term3 * wC * (zij / Rij + zik / Rik);
              /* molecule j */
              A->fx[sj] -= c * term3 * (-fsij * xsij + fsjk * xsjk) * fdmp * wE;
              A->fy[sj] -= c * term3 * (-fsij * ysij + fsjk * ysjk) * fdmp * wE;
              A->fz[sj] -= c * term3 * (-fsij * zsij + fsjk * zsjk) * fdmp * wE;
              A->cx[j] -= c * term3 * wC * (-xij / Rij + xjk / Rjk);
              A->cy[j] -= c * term3 * wC * (-yij / Rij + yjk / Rjk);
              A->cz[j] -= c * term3 * wC * (-zij / Rij + zjk / Rjk);
              /* molecule k */
              A->fx[sk] -= c * term3 * (-fsik * xsik - fsjk * xsjk) * fdmp * wE;
              A->fy[sk] -= c * term3 * (-fsik * ysik - fsjk * ysjk) * fdmp * wE;
              A->fz[sk] -= c * term3 * (-fsik * zsik - fsjk * zsjk) * fdmp * wE;
              A->cx[k] -= c * term3 * wC * (-xik / Rik - xjk / Rjk);
              A->cy[k] -= c * term3 * wC * (-yik / Rik - yjk / Rjk);
              A->cz[k] -= c * term3 * wC * (-zik / Rik - zjk / Rjk);
            }
          }
        }
      }
}

/* The same region sweep as fs3_region, reorganised (not a restatement of the
 * reference's loop nest; used for the optimised CPU baseline and checked
 * against fs3_region by tests/test_oracle_canonical.py).  With the pair vector
 * v_ab = (e_ab, e_ab r_ab), the raw polynomial sum is
 *   E = sum_{a,b,c} sum_{ip,jp,kp} C[a,b,c,ip,jp,kp] v_jk[b,c][ip] v_ik[a,c][jp] v_ij[a,b][kp],
 * so dE/dr of a site pair follows from the gradient G of E with respect to its
 * pair vector: dE/dr = e (G0 der + G1 (r der + 1)), der = d ln e / dr.  Each
 * site pair's functions are evaluated once (867 per region instead of
 * 3 x 4913), and the weights, forces and virial are applied per site pair:
 *   site forces  -(fdmp wE / r) dE/dr x_ab on a, + on b;
 *   virial       fdmp wE r dE/dr per site pair, wC E (R_ij + R_ik + R_jk);
 *   COM forces   -wC E (x_ij / R_ij + x_ik / R_ik) on i, and so on. */
static void fs3_region_canonical(const geom_t *G, int i, int j, int k, const double v[9],
                                 const double R[3], const double (*bS)[17], const double (*gS)[17],
                                 const double (*r0S)[17], const double *cs, const int *tbl,
                                 double fdmp, double wE, double wC, acc_t *A) {
  enum { N = ORC_NS3 };
  const int ld = G->ld;
  const double *xxs = G->xxs, *yys = G->yys, *zzs = G->zzs;
  /* blocks 0: ij [a][b], 1: ik [a][c], 2: jk [b][c] */
  const int mol[3][2] = {{i, j}, {i, k}, {j, k}};
  double sv[3][N][N][3], sr[3][N][N], se[3][N][N][2], sd[3][N][N], sg[3][N][N][2];
  int blk, a, b, c, m;
  for (blk = 0; blk < 3; ++blk) {
    const int m0 = mol[blk][0], m1 = mol[blk][1];
    for (a = 0; a < N; ++a)
      for (b = 0; b < N; ++b) {
        const long s0 = a + (long)ld * m0, s1 = b + (long)ld * m1;
        const double x = v[3 * blk] + (xxs[s0] - xxs[s1]);
        const double y = v[3 * blk + 1] + (yys[s0] - yys[s1]);
        const double z = v[3 * blk + 2] + (zzs[s0] - zzs[s1]);
        const double r = sqrt(x * x + y * y + z * z);
        const double be = bS[a][b], ga = gS[a][b], r0 = r0S[a][b];
        const double t = exp(-ga * (r - r0));
        const double damp = 1.0 / (1.0 + t);
        const double e = exp(-be * r) * damp;
        sv[blk][a][b][0] = x;
        sv[blk][a][b][1] = y;
        sv[blk][a][b][2] = z;
        sr[blk][a][b] = r;
        se[blk][a][b][0] = e;
        se[blk][a][b][1] = e * r;
        sd[blk][a][b] = -be + ga * t * damp;
        sg[blk][a][b][0] = sg[blk][a][b][1] = 0.0;
      }
  }
  double E = 0.0;
  for (a = 0; a < N; ++a)
    for (b = 0; b < N; ++b) {
      const double *zv = se[0][a][b];
      double gij0 = 0.0, gij1 = 0.0;
      for (c = 0; c < N; ++c) {
        const int *tb = tbl + ((a * N + b) * N + c) * 8;
        double C8[8];
        for (m = 0; m < 8; ++m) C8[m] = cs[tb[m] - 1];  /* index ip*4 + jp*2 + kp */
        const double *xv = se[2][b][c], *yv = se[1][a][c];
        const double u00 = xv[0] * yv[0], u01 = xv[0] * yv[1], u10 = xv[1] * yv[0], u11 = xv[1] * yv[1];
        gij0 += C8[0] * u00 + C8[2] * u01 + C8[4] * u10 + C8[6] * u11;
        gij1 += C8[1] * u00 + C8[3] * u01 + C8[5] * u10 + C8[7] * u11;
        /* d/dv_ik[jp]: sum over ip, kp of C x[ip] z[kp] */
        const double w00 = xv[0] * zv[0], w01 = xv[0] * zv[1], w10 = xv[1] * zv[0], w11 = xv[1] * zv[1];
        sg[1][a][c][0] += C8[0] * w00 + C8[1] * w01 + C8[4] * w10 + C8[5] * w11;
        sg[1][a][c][1] += C8[2] * w00 + C8[3] * w01 + C8[6] * w10 + C8[7] * w11;
        /* d/dv_jk[ip]: sum over jp, kp of C y[jp] z[kp] */
        const double q00 = yv[0] * zv[0], q01 = yv[0] * zv[1], q10 = yv[1] * zv[0], q11 = yv[1] * zv[1];
        sg[2][b][c][0] += C8[0] * q00 + C8[1] * q01 + C8[2] * q10 + C8[3] * q11;
        sg[2][b][c][1] += C8[4] * q00 + C8[5] * q01 + C8[6] * q10 + C8[7] * q11;
      }
      sg[0][a][b][0] = gij0;
      sg[0][a][b][1] = gij1;
      E += zv[0] * gij0 + zv[1] * gij1;
    }
  const double fw = fdmp * wE;
  A->e = A->e + fw * E;
  double W = 0.0;
  for (blk = 0; blk < 3; ++blk) {
    const int m0 = mol[blk][0], m1 = mol[blk][1];
    for (a = 0; a < N; ++a)
      for (b = 0; b < N; ++b) {
        const double r = sr[blk][a][b], e = se[blk][a][b][0], der = sd[blk][a][b];
        const double dEdr = e * (sg[blk][a][b][0] * der + sg[blk][a][b][1] * (r * der + 1.0));
        const double g = fw * dEdr / r;
        const long s0 = a + (long)ld * m0, s1 = b + (long)ld * m1;
        const double *x = sv[blk][a][b];
        A->fx[s0] -= g * x[0];
        A->fy[s0] -= g * x[1];
        A->fz[s0] -= g * x[2];
        A->fx[s1] += g * x[0];
        A->fy[s1] += g * x[1];
        A->fz[s1] += g * x[2];
        W += fw * r * dEdr;
      }
  }
  W += wC * E * (R[0] + R[1] + R[2]);
  A->w = A->w + W;
  const double uij[3] = {v[0] / R[0], v[1] / R[0], v[2] / R[0]};
  const double uik[3] = {v[3] / R[1], v[4] / R[1], v[5] / R[1]};
  const double ujk[3] = {v[6] / R[2], v[7] / R[2], v[8] / R[2]};
  A->cx[i] -= wC * E * (uij[0] + uik[0]);
  A->cy[i] -= wC * E * (uij[1] + uik[1]);
  A->cz[i] -= wC * E * (uij[2] + uik[2]);
  A->cx[j] -= wC * E * (-uij[0] + ujk[0]);
  A->cy[j] -= wC * E * (-uij[1] + ujk[1]);
  A->cz[j] -= wC * E * (-uij[2] + ujk[2]);
  A->cx[k] -= wC * E * (-uik[0] - ujk[0]);
  A->cy[k] -= wC * E * (-uik[1] - ujk[1]);
  A->cz[k] -= wC * E * (-uik[2] - ujk[2]);
}

typedef void (*fs3_region_fn)(const geom_t *, int, int, int, const double *, const double *,
                              const double (*)[17], const double (*)[17], const double (*)[17],
                              const double *, const int *, double, double, double, acc_t *);

/* One accepted FS3 triplet: R_avg, damping, region dispatch (:894-1943) */
static int fs3_triplet(const orc_params *P, const geom_t *G, int i, int j, int k,
                       const double v[9], const double d2[3], int *tbl, int rebuild,
                       acc_t *A, fs3_region_fn fs3_region) {
  const double alp3 = P->alpha3 * bohr2a;            /* :771 */
  const double r0dmp = 6.20446 / bohr2a;             /* :772 */
  const double rmin = 4.5 / bohr2a, rmax = 5.5 / bohr2a; /* :774-775 */
  double RAB = sqrt(d2[0]), RBC = sqrt(d2[2]), RAC = sqrt(d2[1]);
  double Ravg = (RAB + RBC + RAC) / 3.0;
  double pomij = exp(alp3 * (Ravg - r0dmp));
  double ppomij = 1.0 / (1.0 + pomij);
  double fdmp = pow(ppomij, P->delta33);
  double dfdmp = -alp3 * P->delta33 * pomij * ppomij * fdmp * (1 / 3.0);
  double R[3];
  R[0] = sqrt(d2[0]);
  R[1] = sqrt(d2[1]);
  R[2] = sqrt(d2[2]);
  if (rebuild) orc_build_ind_lin(tbl);
  if (Ravg <= rmin) {
    fs3_region(G, i, j, k, v, R, P->beta_sr, P->gamma_sr, P->r0_sr, P->cs_sr, tbl, fdmp, 1.0,
               dfdmp * 1.0 + fdmp * 0.0, A);
    return 0;
  } else if (Ravg >= rmax) {
    fs3_region(G, i, j, k, v, R, P->beta_lr, P->gamma_lr, P->r0_lr, P->cs_lr, tbl, fdmp, 1.0,
               dfdmp * 1.0 + fdmp * 0.0, A);
    return 2;
  } else {
    /* cubic switch (:1420-1423) */
    double cfact = (Ravg - rmin) / (rmax - rmin);
    double smfun = 1.0 + 2.0 * pow(cfact, 3.0) - 3.0 * pow(cfact, 2.0);
    double smfun2 = 1.0 - smfun;
    double cf = 2.0 / (rmax - rmin) * (pow(cfact, 2.0) - cfact);
    fs3_region(G, i, j, k, v, R, P->beta_sr, P->gamma_sr, P->r0_sr, P->cs_sr, tbl, fdmp, smfun,
               dfdmp * smfun + fdmp * cf, A);
    fs3_region(G, i, j, k, v, R, P->beta_lr, P->gamma_lr, P->r0_lr, P->cs_lr, tbl, fdmp, smfun2,
               dfdmp * smfun2 - fdmp * cf, A);
    return 1;
  }
}

/* One accepted FS2 item (k; i<j): the ks x is x js body (:339-539) */
static void fs2_item(const orc_params *P, const geom_t *G, int k, int i, int j,
                     const double v[9], const double d2[3], acc_t *A) {
  const int ld = G->ld;
  const double alp3 = P->alpha3 * bohr2a;  /* :189 */
  const double r0dmp = 7.40446 / bohr2a;   /* :191 */
  const double delta33 = P->delta33;
  const double xij = v[0], yij = v[1], zij = v[2];
  const double xki = v[3], yki = v[4], zki = v[5];
  const double xkj = v[6], ykj = v[7], zkj = v[8];
  const double dij2 = d2[0], dki2 = d2[1], dkj2 = d2[2];
  const double *xxs = G->xxs, *yys = G->yys, *zzs = G->zzs;
  double pom, pomki, pomkj, flrki, flrkj, dflrki, dflrkj, dflrkiv, dflrkjv;
  int ks, is, js;

  pom = exp(alp3 * (sqrt(dki2) - r0dmp));
  pomki = 1.0 / (1.0 + pom);
  flrki = pow(pomki, delta33);
  dflrki = -alp3 * delta33 * pom * pomki * flrki;
  dflrkiv = -alp3 * delta33 * pom * pomki * sqrt(dki2);

  pom = exp(alp3 * (sqrt(dkj2) - r0dmp));
  pomkj = 1.0 / (1.0 + pom);
  flrkj = pow(pomkj, delta33);
  dflrkj = -alp3 * delta33 * pom * pomkj * flrkj;
  dflrkjv = -alp3 * delta33 * pom * pomkj * sqrt(dkj2);

  for (ks = 0; ks < 5; ++ks) {
    const double qks = charge3b[ks];
    const long sk = ks + (long)ld * k;
    for (is = 0; is < 3; ++is) {
      for (js = 0; js < 3; ++js) {
        const long si = is + (long)ld * i, sj = js + (long)ld * j;
        double xsij = xij + (xxs[si] - xxs[sj]);
        double ysij = yij + (yys[si] - yys[sj]);
        double zsij = zij + (zzs[si] - zzs[sj]);
        double dsij2 = xsij * xsij + ysij * ysij + zsij * zsij;
        double dsij = sqrt(dsij2);
        double dsij1i = 1.0 / dsij;
        double eqij = eta3 * exp(-beta3 * dsij);

        double dij = sqrt(dij2), dij1i = 1.0 / dij, dij3i = dij1i * dij1i * dij1i;
        double dki = sqrt(dki2), dki1i = 1.0 / dki;
        double dkj = sqrt(dkj2), dkj1i = 1.0 / dkj;

        double dxkqi = xki + xxs[sk] - g3 * xij * dij1i;
        double dykqi = yki + yys[sk] - g3 * yij * dij1i;
        double dzkqi = zki + zzs[sk] - g3 * zij * dij1i;
        double dkqi2 = dxkqi * dxkqi + dykqi * dykqi + dzkqi * dzkqi;
        double dkqi = sqrt(dkqi2), dkqi1i = 1.0 / dkqi;
        double dkqi3i = dkqi1i * dkqi1i * dkqi1i;

        double dxkqj = xkj + xxs[sk] + g3 * xij * dij1i;
        double dykqj = ykj + yys[sk] + g3 * yij * dij1i;
        double dzkqj = zkj + zzs[sk] + g3 * zij * dij1i;
        double dkqj2 = dxkqj * dxkqj + dykqj * dykqj + dzkqj * dzkqj;
        double dkqj = sqrt(dkqj2), dkqj1i = 1.0 / dkqj;
        double dkqj3i = dkqj1i * dkqj1i * dkqj1i;

        double dxkci = xki + xxs[sk], dykci = yki + yys[sk], dzkci = zki + zzs[sk];
        double dkci2 = dxkci * dxkci + dykci * dykci + dzkci * dzkci;
        double dkci = sqrt(dkci2), dkci1i = 1.0 / dkci;
        double dkci3i = dkci1i * dkci1i * dkci1i;

        double dxkcj = xkj + xxs[sk], dykcj = ykj + yys[sk], dzkcj = zkj + zzs[sk];
        double dkcj2 = dxkcj * dxkcj + dykcj * dykcj + dzkcj * dzkcj;
        double dkcj = sqrt(dkcj2), dkcj1i = 1.0 / dkcj;
        double dkcj3i = dkcj1i * dkcj1i * dkcj1i;

        double term0 = (dkci1i + dkcj1i - dkqi1i - dkqj1i);
        double termxy1, termxy2, dotkqi_ij, dotkqj_ij;
        double fxk, fyk, fzk, fxi1, fyi1, fzi1, fxi2, fyi2, fzi2, fxi3, fyi3, fzi3;
        double fxi4, fyi4, fzi4, fxj1, fyj1, fzj1, fxj2, fyj2, fzj2, fxj3, fyj3, fzj3;
        double fxj4, fyj4, fzj4;
        (void)dki1i; (void)dkj1i;

        A->e0 = A->e0 + qks * eqij * term0;
        A->e = A->e + qks * eqij * term0 * flrki * flrkj;

        termxy1 = (dxkqi * xij + dykqi * yij + dzkqi * zij) * g3 * dij1i * dkqi3i;
        termxy2 = (dxkqj * xij + dykqj * yij + dzkqj * zij) * g3 * dij1i * dkqj3i;
        A->w = A->w + ((-beta3 * dsij - 1.0) + dflrkiv + dflrkjv) * qks * eqij * term0 * flrki * flrkj +
               qks * eqij * (termxy1 - termxy2) * flrki * flrkj;

        dotkqi_ij = xij * dxkqi + yij * dykqi + zij * dzkqi;
        dotkqj_ij = xij * dxkqj + yij * dykqj + zij * dzkqj;

        fxk = qks * eqij * (-dxkci * dkci3i - dxkcj * dkcj3i + dxkqi * dkqi3i + dxkqj * dkqj3i);
        fyk = qks * eqij * (-dykci * dkci3i - dykcj * dkcj3i + dykqi * dkqi3i + dykqj * dkqj3i);
        fzk = qks * eqij * (-dzkci * dkci3i - dzkcj * dkcj3i + dzkqi * dkqi3i + dzkqj * dkqj3i);

        A->cx[k] = A->cx[k] - qks * eqij * term0 * dflrki * xki * flrkj / sqrt(dki2) -
                   qks * eqij * term0 * dflrkj * xkj * flrki / sqrt(dkj2);
        A->cy[k] = A->cy[k] - qks * eqij * term0 * dflrki * yki * flrkj / sqrt(dki2) -
                   qks * eqij * term0 * dflrkj * ykj * flrki / sqrt(dkj2);
        A->cz[k] = A->cz[k] - qks * eqij * term0 * dflrki * zki * flrkj / sqrt(dki2) -
                   qks * eqij * term0 * dflrkj * zkj * flrki / sqrt(dkj2);

        fxi1 = qks * eqij * (-beta3) * term0 * xsij * dsij1i;
        fyi1 = qks * eqij * (-beta3) * term0 * ysij * dsij1i;
        fzi1 = qks * eqij * (-beta3) * term0 * zsij * dsij1i;
        fxi2 = qks * eqij * (dxkci * dkci3i);
        fyi2 = qks * eqij * (dykci * dkci3i);
        fzi2 = qks * eqij * (dzkci * dkci3i);
        fxi3 = qks * eqij * dkqi3i * (-(1.0 + g3 * dij1i) * dxkqi + g3 * dotkqi_ij * xij * dij3i);
        fyi3 = qks * eqij * dkqi3i * (-(1.0 + g3 * dij1i) * dykqi + g3 * dotkqi_ij * yij * dij3i);
        fzi3 = qks * eqij * dkqi3i * (-(1.0 + g3 * dij1i) * dzkqi + g3 * dotkqi_ij * zij * dij3i);
        fxi4 = qks * eqij * dkqj3i * (g3 * dxkqj * dij1i - g3 * dotkqj_ij * xij * dij3i);
        fyi4 = qks * eqij * dkqj3i * (g3 * dykqj * dij1i - g3 * dotkqj_ij * yij * dij3i);
        fzi4 = qks * eqij * dkqj3i * (g3 * dzkqj * dij1i - g3 * dotkqj_ij * zij * dij3i);

        fxj1 = qks * eqij * beta3 * term0 * xsij * dsij1i;
        fyj1 = qks * eqij * beta3 * term0 * ysij * dsij1i;
        fzj1 = qks * eqij * beta3 * term0 * zsij * dsij1i;
        fxj2 = qks * eqij * (dxkcj * dkcj3i);
        fyj2 = qks * eqij * (dykcj * dkcj3i);
        fzj2 = qks * eqij * (dzkcj * dkcj3i);
        fxj3 = qks * eqij * dkqi3i * (g3 * dxkqi * dij1i - g3 * dotkqi_ij * xij * dij3i);
        fyj3 = qks * eqij * dkqi3i * (g3 * dykqi * dij1i - g3 * dotkqi_ij * yij * dij3i);
        fzj3 = qks * eqij * dkqi3i * (g3 * dzkqi * dij1i - g3 * dotkqi_ij * zij * dij3i);
        fxj4 = qks * eqij * dkqj3i * (-(1.0 + g3 * dij1i) * dxkqj + g3 * dotkqj_ij * xij * dij3i);
        fyj4 = qks * eqij * dkqj3i * (-(1.0 + g3 * dij1i) * dykqj + g3 * dotkqj_ij * yij * dij3i);
        fzj4 = qks * eqij * dkqj3i * (-(1.0 + g3 * dij1i) * dzkqj + g3 * dotkqj_ij * zij * dij3i);

        A->fx[sk] = A->fx[sk] - fxk * flrki * flrkj;
        A->fy[sk] = A->fy[sk] - fyk * flrki * flrkj;
        A->fz[sk] = A->fz[sk] - fzk * flrki * flrkj;
        A->fx[si] = A->fx[si] - fxi1 * flrki * flrkj;
        A->fy[si] = A->fy[si] - fyi1 * flrki * flrkj;
        A->fz[si] = A->fz[si] - fzi1 * flrki * flrkj;
        A->fx[sj] = A->fx[sj] - fxj1 * flrki * flrkj;
        A->fy[sj] = A->fy[sj] - fyj1 * flrki * flrkj;
        A->fz[sj] = A->fz[sj] - fzj1 * flrki * flrkj;

        A->cx[i] = A->cx[i] - (fxi2 + fxi3 + fxi4) * flrki * flrkj +
                   qks * eqij * term0 * dflrki * xki * flrkj / sqrt(dki2);
        A->cy[i] = A->cy[i] - (fyi2 + fyi3 + fyi4) * flrki * flrkj +
                   qks * eqij * term0 * dflrki * yki * flrkj / sqrt(dki2);
        A->cz[i] = A->cz[i] - (fzi2 + fzi3 + fzi4) * flrki * flrkj +
                   qks * eqij * term0 * dflrki * zki * flrkj / sqrt(dki2);
        A->cx[j] = A->cx[j] - (fxj2 + fxj3 + fxj4) * flrki * flrkj +
                   qks * eqij * term0 * flrki * dflrkj * xkj / sqrt(dkj2);
        A->cy[j] = A->cy[j] - (fyj2 + fyj3 + fyj4) * flrki * flrkj +
                   qks * eqij * term0 * flrki * dflrkj * ykj / sqrt(dkj2);
        A->cz[j] = A->cz[j] - (fzj2 + fzj3 + fzj4) * flrki * flrkj +
                   qks * eqij * term0 * flrki * dflrkj * zkj / sqrt(dkj2);
      }
    }
  }
}

/* Merge worker accumulators into the caller's arrays with the reference's
 * unit conversion (:574-585 / :1964-1977, then :604/:1996, :669-671/:2058). */
static void finish(int nm, int ld, int nsites, acc_t *A, int nacc, double *qout,
                   double *fxs, double *fys, double *fzs, double *fcx, double *fcy, double *fcz,
                   double *vir) {
  const double conv = h2kcal * 418.4 / bohr2a;
  int a, i, is;
  double e = 0.0, w = 0.0;
  for (i = 0; i < nm; ++i) {
    double cx = 0.0, cy = 0.0, cz = 0.0;
    for (a = 0; a < nacc; ++a) {
      cx = cx + A[a].cx[i] * h2kcal * 418.4 / bohr2a;
      cy = cy + A[a].cy[i] * h2kcal * 418.4 / bohr2a;
      cz = cz + A[a].cz[i] * h2kcal * 418.4 / bohr2a;
    }
    fcx[i] = cx;
    fcy[i] = cy;
    fcz[i] = cz;
    for (is = 0; is < nsites; ++is) {
      long s = is + (long)ld * i;
      double sx = 0.0, sy = 0.0, sz = 0.0;
      for (a = 0; a < nacc; ++a) {
        sx = sx + A[a].fx[s] * h2kcal * 418.4 / bohr2a;
        sy = sy + A[a].fy[s] * h2kcal * 418.4 / bohr2a;
        sz = sz + A[a].fz[s] * h2kcal * 418.4 / bohr2a;
      }
      fxs[s] += sx;
      fys[s] += sy;
      fzs[s] += sz;
    }
  }
  (void)conv;
  for (a = 0; a < nacc; ++a) {
    e += A[a].e;
    w += A[a].w;
  }
  *qout = e * h2kcal;
  *vir += w * h2kcal * 418.4;
}

static int nthreads_of(const orc_opts *o) {
  int n = (o && o->nthreads > 1) ? o->nthreads : 1;
#ifndef _OPENMP
  n = 1;
#endif
  return n;
}

void orc_fs3_term(const orc_params *P, int nm, int ld,
                  const double *x, const double *y, const double *z,
                  const double *xs, const double *ys, const double *zs, double box,
                  double *qfs3, double *fxs, double *fys, double *fzs,
                  double *fxfs3, double *fyfs3, double *fzfs3, double *vir,
                  const orc_opts *opts, orc_stats *stats) {
  orc_opts dflt;
  geom_t G;
  bohr_buf B;
  const double rc2 = orc_fs3_rc2_bohr2();
  long n, t, first, last;
  int *list, nt, a;
  acc_t *A;
  long hist[3] = {0, 0, 0};
  if (!opts) {
    orc_default_opts(&dflt);
    opts = &dflt;
  }
  double t0, t1, t2;
  make_geom(nm, ld, ORC_NS3, x, y, z, xs, ys, zs, box, &G, &B);
  t0 = wall_seconds();
  {
    long (*build)(const geom_t *, double, int *, long) = opts->optimised ? fs3_list_nb : fs3_list_bohr;
    n = build(&G, rc2, NULL, 0);
    list = (int *)malloc(sizeof(int) * 3 * (n ? n : 1));
    build(&G, rc2, list, n);
  }
  t1 = wall_seconds();
  first = opts->fs3_first < 0 ? 0 : opts->fs3_first;
  last = (opts->fs3_last < 0 || opts->fs3_last > n) ? n : opts->fs3_last;
  if (first > last) first = last;
  nt = nthreads_of(opts);
  A = (acc_t *)calloc(nt, sizeof(acc_t));
  for (a = 0; a < nt; ++a) acc_alloc(&A[a], nm, ld);

#pragma omp parallel num_threads(nt) if (nt > 1)
  {
    int tid = 0;
    int tbl[ORC_NZERO];
    long lh[3] = {0, 0, 0};
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    orc_build_ind_lin(tbl);
#pragma omp for schedule(dynamic, 1)
    for (t = first; t < last; ++t) {
      double v[9], d2[3];
      int i = list[3 * t], j = list[3 * t + 1], k = list[3 * t + 2];
      fs3_accept(&G, i, j, k, rc2, v, d2);
      lh[fs3_triplet(P, &G, i, j, k, v, d2, tbl, opts->rebuild_table_per_triplet, &A[tid],
                     opts->optimised ? fs3_region_canonical : fs3_region)]++;
    }
#pragma omp critical
    {
      hist[0] += lh[0];
      hist[1] += lh[1];
      hist[2] += lh[2];
    }
  }
  t2 = wall_seconds();
  finish(nm, ld, ORC_NS3, A, nt, qfs3, fxs, fys, fzs, fxfs3, fyfs3, fzfs3, vir);
  if (stats) {
    stats->t_fs3_list = t1 - t0;
    stats->t_fs3_eval = t2 - t1;
    stats->n_fs3 = n;
    stats->n_fs3_sr = hist[0];
    stats->n_fs3_sw = hist[1];
    stats->n_fs3_lr = hist[2];
  }
  for (a = 0; a < nt; ++a) acc_free(&A[a]);
  free(A);
  free(list);
  free_bohr(&B);
}

void orc_fs2_term(const orc_params *P, int nm, int ld,
                  const double *x, const double *y, const double *z,
                  const double *xs, const double *ys, const double *zs, double box,
                  double *qfs2, double *fxs, double *fys, double *fzs,
                  double *fxfs2, double *fyfs2, double *fzfs2, double *vir,
                  const orc_opts *opts, orc_stats *stats) {
  orc_opts dflt;
  geom_t G;
  bohr_buf B;
  const double rc2 = orc_fs2_rc2_bohr2(P->rcut3);
  long n, t, first, last;
  int *list, nt, a;
  acc_t *A;
  if (!opts) {
    orc_default_opts(&dflt);
    opts = &dflt;
  }
  double t0, t1, t2;
  make_geom(nm, ld, 5, x, y, z, xs, ys, zs, box, &G, &B);
  t0 = wall_seconds();
  {
    long (*build)(const geom_t *, double, int *, long) = opts->optimised ? fs2_list_nb : fs2_list_bohr;
    n = build(&G, rc2, NULL, 0);
    list = (int *)malloc(sizeof(int) * 3 * (n ? n : 1));
    build(&G, rc2, list, n);
  }
  t1 = wall_seconds();
  first = opts->fs2_first < 0 ? 0 : opts->fs2_first;
  last = (opts->fs2_last < 0 || opts->fs2_last > n) ? n : opts->fs2_last;
  if (first > last) first = last;
  nt = nthreads_of(opts);
  A = (acc_t *)calloc(nt, sizeof(acc_t));
  for (a = 0; a < nt; ++a) acc_alloc(&A[a], nm, ld);

#pragma omp parallel num_threads(nt) if (nt > 1)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
#pragma omp for schedule(dynamic, 64)
    for (t = first; t < last; ++t) {
      double v[9], d2[3];
      int k = list[3 * t], i = list[3 * t + 1], j = list[3 * t + 2];
      fs2_accept(&G, k, i, j, rc2, v, d2);
      fs2_item(P, &G, k, i, j, v, d2, &A[tid]);
    }
  }
  t2 = wall_seconds();
  finish(nm, ld, 5, A, nt, qfs2, fxs, fys, fzs, fxfs2, fyfs2, fzfs2, vir);
  if (stats) {
    stats->n_fs2 = n;
    stats->t_fs2_list = t1 - t0;
    stats->t_fs2_eval = t2 - t1;
  }
  for (a = 0; a < nt; ++a) acc_free(&A[a]);
  free(A);
  free(list);
  free_bohr(&B);
}

void orc_threebody_term(const orc_params *P, int nm, int ld,
                        const double *x, const double *y, const double *z,
                        const double *xs, const double *ys, const double *zs,
                        double box,
                        double *qfs2, double *qfs3,
                        double *fxs, double *fys, double *fzs,
                        double *fxfs2, double *fyfs2, double *fzfs2,
                        double *fxfs3, double *fyfs3, double *fzfs3,
                        double *vir, const orc_opts *opts, orc_stats *stats) {
  /* :68-80 -- dispatch on the S2_term / S3_term switches */
  if (stats) memset(stats, 0, sizeof(*stats));
  if (P->s2_term)
    orc_fs2_term(P, nm, ld, x, y, z, xs, ys, zs, box, qfs2, fxs, fys, fzs, fxfs2, fyfs2,
                 fzfs2, vir, opts, stats);
  if (P->s3_term)
    orc_fs3_term(P, nm, ld, x, y, z, xs, ys, zs, box, qfs3, fxs, fys, fzs, fxfs3, fyfs3,
                 fzfs3, vir, opts, stats);
}
