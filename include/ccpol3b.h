/*
 * ccpol3b.h -- C ABI of the MI355X three-body library (libccpol3b.so).
 *
 * Drop-in for the reference routine
 *     subroutine threebody_term(xxx,yyy,zzz,xxxs,yyys,zzzs,boxii,qfs2,qfs3)
 * of source/threebody_potential.f90:17-99 together with the module globals it
 * reads (parameters, common_arrays.f90:201-218) and writes (site forces fxs..,
 * COM forces fxfs2../fxfs3.., torques txfs2../txfs3.., virial `vir`).
 *
 * Conventions (identical to the reference):
 *   - positions in Angstrom; site offsets are Fortran column-major arrays with
 *     leading dimension ld_site: xs(is,i) == xs[(is-1) + ld_site*(i-1)]
 *     (common_arrays.f90:22, nsite=50);
 *   - site forces fxs,fys,fzs are ACCUMULATED (+=) at sites 1..17 (FS3) and
 *     1..5 (FS2); fxfs2.. and fxfs3.. are OVERWRITTEN; torque arrays, when given,
 *     are set to zero; `vir` is ACCUMULATED;
 *   - energies in kcal/mol, forces in 10 J mol^-1 A^-1, virial in 10 J mol^-1
 *     (threebody_potential.f90:574-585, 604, 669-671, 1964-1977, 1996, 2058-2059);
 *   - box <= 0 selects an open boundary (isolated clusters).
 * Every entry point returns 0 on success and a negative CCPOL3B_E* code on
 * failure; on failure the caller's output arrays are left untouched.
 * Forces are summed in 128-bit fixed point (units of 2^-96 hartree/bohr), so
 * they are the same bits run to run and whatever the batch a box is evaluated
 * in; a force contribution of 2^30 hartree/bohr or more, or not a number
 * (overlapping sites), makes the call fail with CCPOL3B_EINVAL instead.
 * One context is used by one host thread at a time.
 */
#ifndef CCPOL3B_H
#define CCPOL3B_H

#ifdef __cplusplus
extern "C" {
#endif

#define CCPOL3B_NS3 17
#define CCPOL3B_NLIN 364

enum ccpol3b_precision {
  CCPOL3B_FP64 = 0,  /* everything in FP64 */
  CCPOL3B_MIXED = 1  /* FP32 exponential/polynomial work, FP64 displacements,
                        switch weights and final reductions */
};

enum ccpol3b_status {
  CCPOL3B_OK = 0,
  CCPOL3B_EINVAL = -1,    /* bad argument */
  CCPOL3B_EHIP = -2,      /* a HIP runtime call or kernel launch failed */
  CCPOL3B_ENOMEM = -3,    /* device allocation failed */
  CCPOL3B_ECAPACITY = -4, /* a molecule has more neighbours than the list capacity */
  CCPOL3B_ENODEV = -5     /* no usable gfx950 device */
};

/* Everything threebody_term reads from module globals.
 * Pair tables are Fortran-ordered (is,js): beta(is,js) == beta_sr[(is-1) + ld_pair*(js-1)],
 * so the reference's bS3_s(nsite,nsite) arrays can be passed with ld_pair = 50.
 * alpha3 (1/A), delta33 and rcut3 (A) are CONTROL record 15; s2_term/s3_term record 14. */
typedef struct {
  const double *beta_sr, *gamma_sr, *r0_sr;   /* bS3_s, gS3_s, r0S3_s */
  const double *beta_lr, *gamma_lr, *r0_lr;   /* bS3_l, gS3_l, r0S3_l */
  int ld_pair;                                /* >= 17 */
  const double *cs_sr, *cs_lr;                /* cs3_s(1:364), cs3_l(1:364) */
  int nlin;                                   /* must be 364 */
  double alpha3, delta33, rcut3;
  int s2_term, s3_term;
  int precision;                              /* enum ccpol3b_precision */
  int device;                                 /* HIP device ordinal, -1 = current */
} ccpol3b_params;

typedef struct ccpol3b_ctx ccpol3b_ctx;

/* Counters and timings of the last evaluation. */
typedef struct {
  long n_fs3;                        /* unordered FS3 triplets i<j<k accepted */
  long n_fs3_sr, n_fs3_switch, n_fs3_lr; /* region histogram (R_avg <=4.5, between, >=5.5 A) */
  long n_fs2_items;                  /* ordered FS2 items (k; i<j) accepted */
  long n_fs2_pairs;                  /* unordered FS2 pairs within rcut3 */
  double ms_total;                   /* device time of the whole evaluation (HIP events) */
  double ms_lists;                   /* conversion, neighbour lists and their offsets */
  double ms_fs2, ms_fs3;             /* the two concurrent branches after the lists: the FS2 passes,
                                        and the FS3 list build plus the FS3 evaluation */
  double ms_reduce;                  /* from the later branch's end: write-back and energy sums */
  double ms_h2d, ms_d2h;             /* host-pointer entry point only */
  double ms_fs3_eval;                /* the FS3 evaluation kernel alone (inside ms_fs3) */
  long n_cells;                      /* neighbour-search cells over the call's boxes (0: every box
                                        swept molecule by molecule) */
} ccpol3b_stats;

/* Create / destroy a context holding device copies of the parameters.
 * Replaces the parameter reads of initialization.f90:640-701 (values passed in). */
int ccpol3b_create(const ccpol3b_params *params, ccpol3b_ctx **ctx);
void ccpol3b_destroy(ccpol3b_ctx *ctx);

/* Switch the arithmetic mode of an existing context. */
int ccpol3b_set_precision(ccpol3b_ctx *ctx, int precision);

/* The drop-in: host arrays in, host arrays out, same semantics as
 * threebody_term (threebody_potential.f90:17-99).  Torque pointers may be NULL. */
int ccpol3b_threebody_term(ccpol3b_ctx *ctx, int nm, int ld_site,
                           const double *x, const double *y, const double *z,
                           const double *xs, const double *ys, const double *zs,
                           double box, double *qfs2, double *qfs3,
                           double *fxs, double *fys, double *fzs,
                           double *fxfs2, double *fyfs2, double *fzfs2,
                           double *fxfs3, double *fyfs3, double *fzfs3,
                           double *txfs2, double *tyfs2, double *tzfs2,
                           double *txfs3, double *tyfs3, double *tzfs3,
                           double *vir);

/* Same computation on device-resident arrays (same layouts and units), enqueued
 * on `stream` (a hipStream_t; NULL = default stream).  `energies` receives
 * {qfs2, qfs3, vir_increment} on the device.  Site forces are accumulated,
 * COM forces overwritten, as above.  The call returns after the work is
 * enqueued and the list sizes were checked (one small device-to-host read). */
int ccpol3b_eval_device(ccpol3b_ctx *ctx, int nm, int ld_site,
                        const double *d_x, const double *d_y, const double *d_z,
                        const double *d_xs, const double *d_ys, const double *d_zs,
                        double box,
                        double *d_fxs, double *d_fys, double *d_fzs,
                        double *d_fcom2, double *d_fcom3, /* [3*nm] x..,y..,z.. blocks */
                        double *d_energies, void *stream);

/* Allocation-free variant of ccpol3b_eval_device (same arguments, layouts,
 * units and output semantics).  No counterpart in the reference: it is for
 * callers that cannot allocate device memory -- tight serving loops and
 * harnesses that forbid hipMalloc.  Every per-call device array lives in one
 * fixed 32 MiB device array of the library (one copy per GPU, held until the
 * process ends), so the call makes no hipMalloc, hipFree or host-to-device
 * copy; it reads the list sizes back once (one small device-to-host read) to
 * check them.  Limits: at most 1024 molecules (else CCPOL3B_ECAPACITY before
 * any device work), at most 128 FS3/FS2 neighbours per molecule and 98304 FS3
 * triplets (else CCPOL3B_ECAPACITY, outputs untouched).  All calls on one GPU
 * share that array: calls that overlap on one GPU (several host threads or
 * streams) race.  ccpol3b_get_fs3_list reads its last list;
 * ccpol3b_get_fs2_list is not available after it. */
int ccpol3b_eval_device_noalloc(ccpol3b_ctx *ctx, int nm, int ld_site,
                                const double *d_x, const double *d_y, const double *d_z,
                                const double *d_xs, const double *d_ys, const double *d_zs,
                                double box,
                                double *d_fxs, double *d_fys, double *d_fzs,
                                double *d_fcom2, double *d_fcom3,
                                double *d_energies, void *stream);

/* Batched evaluation of nbox independent boxes in one launch sequence.
 * Molecules of box b are rows mol_start[b] .. mol_start[b+1]-1 of the packed
 * device arrays; box_len[b] (A, host array) <= 0 means open boundary.
 * d_energies receives 3 doubles per box {qfs2, qfs3, vir_increment}. */
int ccpol3b_eval_batch_device(ccpol3b_ctx *ctx, int nbox, const int *mol_start,
                              const double *box_len, int ld_site,
                              const double *d_x, const double *d_y, const double *d_z,
                              const double *d_xs, const double *d_ys, const double *d_zs,
                              double *d_fxs, double *d_fys, double *d_fzs,
                              double *d_fcom2, double *d_fcom3,
                              double *d_energies, void *stream);

/* Gradient path (no counterpart in the reference, which has no derivative with
 * respect to its fitted parameters): the batched evaluation above plus, per
 * box, dE/dc = d(qfs3)/d(cs3_s(1:364), cs3_l(1:364)) in kcal/mol per unit
 * coefficient, written to d_phi[box][728] (SR then LR).  FS2 has no fitted
 * linear coefficients.  E is linear in c, so sum_n d_phi[b][n] c_n = qfs3(b). */
int ccpol3b_eval_batch_grad_device(ccpol3b_ctx *ctx, int nbox, const int *mol_start,
                                   const double *box_len, int ld_site,
                                   const double *d_x, const double *d_y, const double *d_z,
                                   const double *d_xs, const double *d_ys, const double *d_zs,
                                   double *d_fxs, double *d_fys, double *d_fzs,
                                   double *d_fcom2, double *d_fcom3,
                                   double *d_energies, double *d_phi, void *stream);

/* Gram matrix gram[a][b] = sum_r phi[r][a] phi[r][b] (row-major, FP64) of a
 * design matrix of coefficient gradients (rows = configurations), on the FP64
 * matrix cores (the FP64 dot-product error bound holds; partial sums over row
 * ranges are added in a fixed order, so results repeat bit for bit).
 * No counterpart in the reference (its refit code is commented out,
 * threebody_potential.f90:1071,1304).  Enqueued on `stream`; the context owns
 * the scratch. */
int ccpol3b_gram_device(ccpol3b_ctx *ctx, int nrow, int ncol, const double *d_phi, int ld_phi,
                        double *d_gram, int ld_gram, void *stream);

/* Replace the linear SR/LR coefficients cs3_s, cs3_l (364 each) of a context,
 * e.g. during a refit (initialization.f90:660-683 read them once). */
int ccpol3b_set_coefficients(ccpol3b_ctx *ctx, const double *cs_sr, const double *cs_lr);

/* Launch shapes of a context.  No counterpart in the reference (a CPU loop
 * nest); chosen per workload by ccpol3b_autotune or set by hand. */
typedef struct {
  int fs3_lds_coef;           /* 1: FS3 blocks stage the region's coefficient table in LDS;
                                 0: they read it through the vector cache (default) */
  int fs2_parts;              /* FS2 spectator-pass blocks per molecule, 1..8 (default 1) */
  int fs2_pair_parts;         /* FS2 pair-pass blocks per molecule, 1..8 (default 1) */
  int neighbour_cells;        /* neighbour search in periodic boxes: 1 (default) a cell grid where
                                 the box holds at least 6 cells a side of the larger cutoff, 2 a
                                 grid from 4 cells a side, 0 every molecule of the box tested.
                                 The lists are the same either way. */
} ccpol3b_tuning;

int ccpol3b_get_tuning(const ccpol3b_ctx *ctx, ccpol3b_tuning *tuning);
int ccpol3b_set_tuning(ccpol3b_ctx *ctx, const ccpol3b_tuning *tuning);

/* Time every candidate launch shape on the given batch (ccpol3b_eval_batch_device
 * arguments, `reps` timed evaluations each after one warm-up, HIP events on
 * `stream`), keep the fastest in the context and return it.  The context
 * allocates its own output scratch for the trials; the inputs are not modified. */
int ccpol3b_autotune(ccpol3b_ctx *ctx, int nbox, const int *mol_start, const double *box_len, int ld_site,
                     const double *d_x, const double *d_y, const double *d_z,
                     const double *d_xs, const double *d_ys, const double *d_zs,
                     int reps, void *stream, ccpol3b_tuning *best, double *best_ms);

/* A batched evaluation recorded once into a HIP graph and replayed with one
 * launch, for fixed shapes and device pointers (no counterpart in the
 * reference).  Creation runs one checked evaluation on a private copy of the
 * context (so its outputs are valid), enlarges the list capacities by
 * `headroom` (e.g. 0.25 = 25 %) and records the launch sequence.  A replay
 * never waits for the host; after waiting for the stream the caller reads
 * ccpol3b_graph_status, which is CCPOL3B_ECAPACITY when the geometry outgrew
 * the recorded capacities, or CCPOL3B_EINVAL when a force left the
 * fixed-point range (that replay wrote none of the outputs).  Outputs
 * follow ccpol3b_eval_batch_device (site forces accumulated); d_phi may be NULL. */
typedef struct ccpol3b_graph ccpol3b_graph;
int ccpol3b_graph_create(ccpol3b_ctx *ctx, int nbox, const int *mol_start, const double *box_len, int ld_site,
                         const double *d_x, const double *d_y, const double *d_z,
                         const double *d_xs, const double *d_ys, const double *d_zs,
                         double *d_fxs, double *d_fys, double *d_fzs,
                         double *d_fcom2, double *d_fcom3, double *d_energies, double *d_phi,
                         double headroom, void *stream, ccpol3b_graph **graph);
int ccpol3b_graph_launch(ccpol3b_graph *graph, void *stream);
int ccpol3b_graph_status(const ccpol3b_graph *graph);
void ccpol3b_graph_destroy(ccpol3b_graph *graph);

/* Rigid-water NVE driver on the device (replaces the rigid-body steps of
 * nve_1, evolve.f90:621-1062, with NO_SQUISH rotation, :8107-8204, for the
 * three-body forces of this library; the caller's two-body, Ewald and
 * induction terms are not part of it).  pcoord: [ns][3] principal-frame site
 * coordinates (A, initialization.f90:728-845), ns >= 17; mass (amu) and the
 * principal moments pmi[3] (amu A^2); box (A, <= 0 open).  State arrays on
 * the host: com, vel (A/ps) and body-frame angular momentum jbody
 * (amu A^2/ps) [nm][3], quaternions [nm][4].  ccpol3b_md_run advances nsteps
 * steps of dt (ps) and writes, per step, {E3 = qfs2 + qfs3, translational and
 * rotational kinetic energy, their sum} in kcal/mol to records[nsteps][4]
 * (may be NULL).  The driver borrows ctx (precision, parameters). */
typedef struct ccpol3b_md ccpol3b_md;
int ccpol3b_md_create(ccpol3b_ctx *ctx, int nm, double box, int ns, const double *pcoord, double mass,
                      const double *pmi, ccpol3b_md **md);
int ccpol3b_md_set_state(ccpol3b_md *md, const double *com, const double *quat, const double *vel,
                         const double *jbody);
int ccpol3b_md_get_state(ccpol3b_md *md, double *com, double *quat, double *vel, double *jbody);
int ccpol3b_md_run(ccpol3b_md *md, int nsteps, double dt, double *records);
void ccpol3b_md_destroy(ccpol3b_md *md);

int ccpol3b_last_stats(const ccpol3b_ctx *ctx, ccpol3b_stats *stats);

/* Copy the last FS3 triplet list (0-based i<j<k, lexicographic order, the
 * reference loop order of threebody_potential.f90:859-892) or the last FS2
 * item list (rows k,i,j in the reference loop order :297-301) to host memory.
 * Returns the number of entries, or a negative error; writes at most cap rows. */
long ccpol3b_get_fs3_list(ccpol3b_ctx *ctx, int *ijk, long cap);
long ccpol3b_get_fs2_list(ccpol3b_ctx *ctx, int *kij, long cap);

/* Reader for one data3b_full* file (initialization.f90:660-680): the 17x17
 * (beta, gamma, r0) table in Fortran (is,js) order with leading dimension
 * ld_pair, and up to *nlin coefficients (on return *nlin = nlin0 of the file). */
int ccpol3b_read_data3b(const char *path, double *beta, double *gamma, double *r0,
                        int ld_pair, double *cs, int *nlin);

/* The orbit number (1..364) of (site i, site j, site k, power of r_jk,
 * power of r_ik, power of r_ij), 1-based sites 1..17: ind_lin_2 of
 * threebody_potential.f90:909-946 built with fill (:2064-2131). */
int ccpol3b_orbit_index(int is, int js, int ks, int ip, int jp, int kp);

const char *ccpol3b_status_string(int status);
const char *ccpol3b_last_error(const ccpol3b_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
