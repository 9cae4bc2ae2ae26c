/*
 * ccpol3b_oracle.h -- CPU FP64 oracle for the CCpol23+(2025) three-body term.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * routine `threebody_term` (FS2_term + FS3_term) of
 * source/threebody_potential.f90, used by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg as the checker.  The product library
 * (libccpol3b.so) never links, loads or calls it.
 *
 * Conventions follow the reference exactly:
 *   - inputs in Angstrom, internally converted to bohr by division by bohr2a
 *     (threebody_potential.f90:48-62);
 *   - site arrays are Fortran column-major with leading dimension `ld`
 *     (xs(is,i) == xs[(is-1) + ld*(i-1)]), common_arrays.f90:22;
 *   - outputs: site forces fxs.. are ACCUMULATED (+=), COM forces fxfs2/fxfs3
 *     are OVERWRITTEN, torques zeroed, `vir` accumulated, energies in kcal/mol,
 *     forces in 10 J/mol/A, virial in 10 J/mol (threebody_potential.f90:574-671,
 *     1964-2059).
 */
#ifndef CCPOL3B_ORACLE_H
#define CCPOL3B_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NS3 17     /* FS3 sites per molecule (threebody_potential.f90:721) */
#define ORC_NLIN 364   /* linear coefficients per region (data3b_full*:292) */
#define ORC_NZERO 39304 /* 17^3 * 2^3 entries of ind_lin_2 (:722-723) */

typedef struct {
  /* [is][js], 0-based; the reference stores bS3_s(is,js) (common_arrays.f90:208-210) */
  double beta_sr[17][17], gamma_sr[17][17], r0_sr[17][17];
  double beta_lr[17][17], gamma_lr[17][17], r0_lr[17][17];
  double cs_sr[ORC_NLIN], cs_lr[ORC_NLIN];
  double alpha3;   /* CONTROL record 15, 1/A */
  double delta33;  /* CONTROL record 15 */
  double rcut3;    /* CONTROL record 15, A (FS2 cutoff) */
  int s2_term, s3_term; /* CONTROL record 14 */
} orc_params;

typedef struct {
  int rebuild_table_per_triplet; /* reproduce the per-triplet ind_lin_2 rebuild (:909-946) */
  int nthreads;                  /* >1: OpenMP over the outer index (timing only) */
  long fs3_first, fs3_last;      /* evaluate only accepted FS3 triplets with ordinal in
                                    [first,last) (bounded timing sample); last<0 = all */
  long fs2_first, fs2_last;      /* same for FS2 accepted (k;i<j) items */
  int optimised;                 /* 1: the optimised CPU baseline: the same FS3 and FS2 lists
                                    built from neighbour lists (O(N n_nb^2) instead of the
                                    O(N^3) scans), and each FS3 region evaluated by the
                                    factorised contraction (fs3_region_canonical: site-pair
                                    functions once per site pair, the trilinear polynomial
                                    summed over site triples with its three pair gradients)
                                    instead of the reference's loop nest; same lists, same
                                    results to rounding */
} orc_opts;

typedef struct {
  long n_fs3;          /* accepted FS3 triplets (all, regardless of sample window) */
  long n_fs3_sr, n_fs3_sw, n_fs3_lr; /* region histogram */
  long n_fs2;          /* accepted ordered FS2 items (k; i<j) */
  double t_fs3_list, t_fs3_eval; /* wall seconds: O(N^3) count + fill scans, evaluation of the window */
  double t_fs2_list, t_fs2_eval;
} orc_stats;

void orc_default_opts(orc_opts *o);

/* Restatement of the first-encounter orbit numbering built from `fill`
 * (threebody_potential.f90:909-946 with :2064-2131).  ind_lin_2 has ORC_NZERO
 * entries, 0-based array whose values are 1..364 (Fortran numbering). */
int orc_build_ind_lin(int *ind_lin_2);

/* Full reference routine (threebody_potential.f90:17-99). Box <= 0 selects an
 * open boundary (no minimum image), identical to an infinitely large box. */
void orc_threebody_term(const orc_params *P, int nm, int ld,
                        const double *x, const double *y, const double *z,
                        const double *xs, const double *ys, const double *zs,
                        double box,
                        double *qfs2, double *qfs3,
                        double *fxs, double *fys, double *fzs,
                        double *fxfs2, double *fyfs2, double *fzfs2,
                        double *fxfs3, double *fyfs3, double *fzfs3,
                        double *vir, const orc_opts *opts, orc_stats *stats);

/* Separate terms (both take Angstrom inputs like orc_threebody_term). */
void orc_fs2_term(const orc_params *P, int nm, int ld,
                  const double *x, const double *y, const double *z,
                  const double *xs, const double *ys, const double *zs, double box,
                  double *qfs2, double *fxs, double *fys, double *fzs,
                  double *fxfs2, double *fyfs2, double *fzfs2, double *vir,
                  const orc_opts *opts, orc_stats *stats);
void orc_fs3_term(const orc_params *P, int nm, int ld,
                  const double *x, const double *y, const double *z,
                  const double *xs, const double *ys, const double *zs, double box,
                  double *qfs3, double *fxs, double *fys, double *fzs,
                  double *fxfs3, double *fyfs3, double *fzfs3, double *vir,
                  const orc_opts *opts, orc_stats *stats);

/* Triplet lists in the reference's loop order.  FS3: unordered i<j<k
 * (threebody_potential.f90:859-892), written as 0-based (i,j,k) int triples.
 * FS2: ordered (k; i<j) items (:297-336), written as (k,i,j).  Return the
 * number of accepted entries; write at most `cap` entries when out != NULL. */
long orc_fs3_list(const orc_params *P, int nm, const double *x, const double *y,
                  const double *z, double box, int *out, long cap);
long orc_fs2_list(const orc_params *P, int nm, const double *x, const double *y,
                  const double *z, double box, int *out, long cap);

/* Cutoff constants exactly as the reference forms them (FP64, libm pow). */
double orc_fs3_rc2_bohr2(void);
double orc_fs2_rc2_bohr2(double rcut3);

#ifdef __cplusplus
}
#endif
#endif
