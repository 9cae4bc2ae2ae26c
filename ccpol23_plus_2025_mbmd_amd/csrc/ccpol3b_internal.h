// Internal definitions shared by the HIP translation units of libccpol3b.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct ccpol3b_graph;  // include/ccpol3b.h
struct ccpol3b_ctx;

namespace ccpol3b {

constexpr int NS3 = 17;        // FS3 sites per molecule
constexpr int NS2 = 5;         // FS2 spectator sites (O, H, H, D1, D1)
constexpr int NTYPE = 6;       // FS3 site types (O, H, D1, N1, N2, N3)
constexpr int NPAIR = NS3 * NS3;      // 289 site pairs per molecule pair
constexpr int NPAIR3 = 3 * NPAIR;     // 867 site pairs per triplet
constexpr int NCOEF = NTYPE * NTYPE * NTYPE * 8;  // dense coefficients per region
constexpr int NLIN3 = 364;     // linear coefficients per region (orbits)
constexpr int ACC_W = NS3 * 3 + 6;     // per-molecule accumulator: 17 site forces, COM(FS2), COM(FS3)
constexpr int ACC_C2 = NS3 * 3;        // offset of the FS2 COM force
constexpr int ACC_C3 = NS3 * 3 + 3;    // offset of the FS3 COM force

// indices into Dev::totals
constexpr int T_N3 = 0;      // FS3 triplets
constexpr int T_NP2 = 1;     // FS2 pairs
constexpr int T_NI2 = 2;     // FS2 items
constexpr int T_NBOVER = 3;  // largest neighbour count when a list overflowed
constexpr int T_HIST = 4;    // [4..6] FS3 region histogram SR, switch, LR
constexpr int T_SL2 = 8;     // FS2 sub-lists (k; i)
constexpr int T_NI2L = 9;    // FS2 items of an exported list
constexpr int T_NCH3 = 10;   // FS3 / FS2 energy reduction chunks
constexpr int T_NCH2 = 11;
constexpr int T_FXOVER = 12; // a force contribution fell outside the fixed-point range
constexpr int T_DONE = 13;   // epilogue blocks finished (the last one forms the box sums)
constexpr int T_COUNT = 16;
constexpr int FS2_PARTS_MAX = 8; // spectator-pass blocks per molecule, at most
constexpr int RED_CHUNK = 2048;  // entries per block in the first level of the energy reduction

constexpr double BOHR2A = 0.529177249;  // common_arrays.f90:34
constexpr double H2KCAL = 627.510;      // common_arrays.f90:35

// site -> type (0-based), threebody_potential.f90:2076
__host__ __device__ constexpr int site_type(int s) {
  return s == 0 ? 0 : s <= 2 ? 1 : s <= 4 ? 2 : s <= 8 ? 3 : s <= 12 ? 4 : 5;
}

// Interleaved layout of a region's coefficient table, read by the FP64
// evaluation kernel: entry (tu, tw, tz, k) of the dense table
// C[tu][tw][tz][k] (k = 4 p + 2 q + s) sits at cil_index(tz, k) + 2 (tu NTYPE + tw).
// For one type block tz and one pair of coefficients (2 k4, 2 k4 + 1) the 36
// type pairs are adjacent, so a 16-byte load of the lanes of a wavefront, which
// hold different type pairs, touches 576 contiguous bytes instead of 36 lines.
constexpr int CIL_K4 = NTYPE * NTYPE * 2;  // step of k / 2
constexpr int CIL_TZ = 4 * CIL_K4;         // step of tz
__host__ __device__ constexpr int cil_index(int tz, int k) { return tz * CIL_TZ + (k >> 1) * CIL_K4 + (k & 1); }
static_assert(cil_index(NTYPE - 1, 7) + 2 * (NTYPE * NTYPE - 1) == NCOEF - 1, "interleaved layout spans the table");

// Mixed-precision FS3 evaluation: the second component of each site-pair
// vector is e (r - MIXED_VSHIFT) instead of e r, with the coefficient table
// transformed to that basis exactly (in FP64, once per coefficient upload).
// The energy is the same polynomial; the shifted basis cancels 6-10x less
// (sum of |terms| / |E| per triplet), and so loses that much less to FP32
// rounding of the terms.  6.5 bohr is the typical site-site distance of the
// triplets; the optimum of either shipped parameter set lies within 6-7 bohr.
constexpr double MIXED_VSHIFT = 6.5;

// Constants of one evaluation (bohr / hartree units).
struct Consts {
  double alp3;     // alpha3 * bohr2a (1/bohr)
  double delta33;
  int idelta;      // delta33 when it is a whole number in [1, 64], else 0
  double r0dmp3;   // 6.20446 A in bohr (FS3 damping centre, :772)
  double r0dmp2;   // 7.40446 A in bohr (FS2 damping centre, :191)
  double rc2_3;    // FS3 COM cutoff^2 (bohr^2), :770
  double rc2_2;    // FS2 COM cutoff^2 (bohr^2), :190
  double rmin, rmax;  // switch window in bohr, :774-775
};

// Geometry of one FS3 triplet, computed once per triplet by the lane of the
// list fill that emits it, and read by every thread of its evaluation block.
struct TriGeo {
  double c[3][3];      // COM vectors x_ij, x_ik, x_jk (bohr), as the reference forms them (:862-884)
  double R[3];         // |x_ij|, |x_ik|, |x_jk|
  double Ravg, fd, dfd;
  double w[2], wp[2];  // region weights (SR, LR) and their R_avg derivatives
  int nreg, reg0, m[3];
  int pad;
};

// Smallest cell grid used: with 3 cells a side the 27-cell stencil is the
// whole box, so the sweep over the box does the same with less work.  By
// default a grid is used from CELL_AUTO cells a side, where the measured list
// build gets faster than the sweep.
constexpr int CELL_MIN = 4;
constexpr int CELL_AUTO = 6;

// Device arrays of one evaluation (sized by the context; see api.hip).
struct Dev {
  int nm, nbox, cap_nb;
  int fs2_parts;              // spectator-pass blocks per molecule (1..FS2_PARTS_MAX)
  int fs2_pair_parts;         // pair-pass blocks per molecule (1..FS2_PARTS_MAX)
  long cap3, cap_pairs2, cap_items2, cap_sl2;
  long grid3;                 // blocks of the FS3 evaluation (<= cap3): a longer list overflows like cap3
  // geometry in bohr
  double *cx, *cy, *cz;   // [nm]
  double *site;           // [nm][17][3] site offsets (bohr)
  int *mol_box;           // [nm]
  double *boxL;           // [nbox] (bohr, <= 0 open)
  int *box_start;         // [nbox+1]
  // cell grids of the periodic boxes wide enough for one (nc >= CELL_MIN cells
  // a side, each side at least the larger cutoff); the other boxes are swept
  int ncells;             // cells of all boxes
  int *cell_nc;           // [nbox] cells a side, 0: no grid
  int *cell0;             // [nbox] first cell of the box
  int *cell_cnt;          // [ncells] molecules per cell (counted by k_prep, back to 0 after the fill)
  int *cell_start;        // [ncells+1] first entry of each cell in cell_mol
  int *cell_mol;          // [nm] molecules grouped by cell
  int *mol_cell;          // [nm] cell of each molecule (-1: no grid)
  // neighbour lists (ascending molecule index) and their first upper entries
  int *nb3_cnt, *nb3, *up3;   // FS3 cutoff: [nm], [nm][cap_nb], [nm]
  int *nb2_cnt, *nb2, *up2;   // FS2 cutoff
  double *nbv2;               // [nm][3][cap_nb] x_i - x_nb (min image, bohr) of each FS2 entry
  // FS2 sub-list (k; i) counts/offsets
  int *sl2_cnt, *sl2_off;     // [nm], [nm+1]
  int *cnt3s, *off3s;         // [nm], [nm+1] FS3 triplets per molecule i (first of the triplet)
  int *cnt2s, *off2s;         // [cap_sl2], [cap_sl2+1] items per FS2 sub-list
  // FS3 triplets (i<j<k, lexicographic)
  int4 *tri;                  // [cap3] (i,j,k,-)
  TriGeo *geo3;               // [cap3] geometry of each triplet
  // FS2 pairs (i<j within the FS2 cutoff): pair index poff2[i] + q - up2[i] for
  // j = nb2(i)[q]; k-major items (k,i,j,pair), built only for list export
  int *pcnt2, *poff2;         // [nm], [nm+1]
  int4 *items2;               // [cap_items2]
  // results (hartree, bohr)
  unsigned long long *acc;    // [nm][ACC_W][2] force accumulators, 128-bit fixed point (devcommon.h)
  double *ew3;                // [cap3][2] per-triplet energy, virial
  double *ew2;                // [nm][fs2_parts][2] FS2 energy, virial per spectator-pass block
  int *box_t3, *box_t2;       // [nbox+1] triplet / FS2 energy-entry ranges of each box
  int *box_c3, *box_c2;       // [nbox+1] first reduction chunk of each box
  double *red3, *red2;        // [chunks][2] first-level energy/virial partial sums
  long cap_red;               // capacity of red3/red2 in chunks
  double *s2pair;             // [cap_pairs2][2] S_ij and dS/dlambda of each FS2 pair
  int *totals;                // [T_COUNT]
};

// The device counters of a graph's private context (totals[T_COUNT], zeroed and
// refilled by every replay) and the triplet capacity it was recorded with: a
// replay overflowed when totals[T_NBOVER] > 0 or totals[T_N3] > cap3.
const int* graph_device_totals(const ::ccpol3b_graph* g, long* cap3);

// A counter that changes whenever the context's precision, coefficients or
// tuning change (what a recorded graph holds a private copy of).
long context_version(const ::ccpol3b_ctx* c);

}  // namespace ccpol3b
