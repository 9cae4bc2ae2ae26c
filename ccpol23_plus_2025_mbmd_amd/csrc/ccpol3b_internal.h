// Internal definitions shared by the HIP translation units of libccpol3b.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ccpol3b {

constexpr int NS3 = 17;        // FS3 sites per molecule
constexpr int NS2 = 5;         // FS2 spectator sites (O, H, H, D1, D1)
constexpr int NTYPE = 6;       // FS3 site types (O, H, D1, N1, N2, N3)
constexpr int NPAIR = NS3 * NS3;      // 289 site pairs per molecule pair
constexpr int NPAIR3 = 3 * NPAIR;     // 867 site pairs per triplet
constexpr int NCOEF = NTYPE * NTYPE * NTYPE * 8;  // dense coefficients per region
constexpr int SLOT3 = NS3 * 3 + 3;    // FS3 per-role output: 17 site forces + COM force
constexpr int PSLOT2 = 2 + 2 * (3 * 3 + 3);  // FS2 per-pair slot: S, WT, i:(3 sites+COM), j:(...)
constexpr int SPEC2 = NS2 * 3 + 3 + 2;       // FS2 per-spectator: 5 site forces, COM, E, W

constexpr double BOHR2A = 0.529177249;  // common_arrays.f90:34
constexpr double H2KCAL = 627.510;      // common_arrays.f90:35

// site -> type (0-based), threebody_potential.f90:2076
__host__ __device__ constexpr int site_type(int s) {
  return s == 0 ? 0 : s <= 2 ? 1 : s <= 4 ? 2 : s <= 8 ? 3 : s <= 12 ? 4 : 5;
}

// Constants of one evaluation (bohr / hartree units).
struct Consts {
  double alp3;     // alpha3 * bohr2a (1/bohr)
  double delta33;
  double r0dmp3;   // 6.20446 A in bohr (FS3 damping centre, :772)
  double r0dmp2;   // 7.40446 A in bohr (FS2 damping centre, :191)
  double rc2_3;    // FS3 COM cutoff^2 (bohr^2), :770
  double rc2_2;    // FS2 COM cutoff^2 (bohr^2), :190
  double rmin, rmax;  // switch window in bohr, :774-775
};

// Device arrays of one evaluation (sized by the context; see api.hip).
struct Dev {
  int nm, nbox, cap_nb;
  long cap3, cap_pairs2, cap_items2;
  // geometry in bohr
  double *cx, *cy, *cz;   // [nm]
  double *site;           // [nm][17][3] site offsets (bohr)
  int *mol_box;           // [nm]
  double *boxL;           // [nbox] (bohr, <= 0 open)
  int *box_start;         // [nbox+1]
  // neighbour lists (ascending molecule index)
  int *nb3_cnt, *nb3;     // FS3 cutoff: [nm], [nm][cap_nb]
  int *nb2_cnt, *nb2;     // FS2 cutoff
  // FS3 triplets (i<j<k, lexicographic)
  int *cnt3, *off3;       // [nm], [nm+1]
  int4 *tri;              // [cap3] (i,j,k,region)
  // FS2 pairs (i<j within rcut3) and k-major items (k,i,j,pair)
  int *pcnt2, *poff2;     // [nm], [nm+1]
  int2 *pairs2;           // [cap_pairs2]
  int *cnt2, *off2;       // [nm], [nm+1]
  int4 *items2;           // [cap_items2]
  // per-unit results (hartree, bohr)
  double *slot3;          // [cap3][3][SLOT3]
  double *ew3;            // [cap3][2] energy, virial
  double *pslot2;         // [cap_pairs2][PSLOT2]
  double *spec2;          // [nm][SPEC2]
  // list totals and flags: [0] n3, [1] npairs2, [2] nitems2, [3] nb overflow,
  // [4..6] region histogram
  int *totals;
};

}  // namespace ccpol3b
