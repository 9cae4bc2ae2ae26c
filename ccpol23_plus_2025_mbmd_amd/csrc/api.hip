// Host side of libccpol3b.so: the C ABI declared in include/ccpol3b.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/ccpol3b.h"
#include "ccpol3b_internal.h"

namespace ccpol3b {
// kernels (lists.hip, eval.hip, fs2.hip)
__global__ void k_prep(int, int, const double*, const double*, const double*, const double*,
                       const double*, const double*, Dev);
__global__ void k_neighbours(Dev, Consts, int);
template <int NT> __global__ void k_cells_build(Dev);
__global__ void k_fs3_fill(Dev, Consts);
__global__ void k_fs2_items(Dev, Consts, int);
template <int NT> __global__ void k_scan(Dev, int, int, int);
__global__ void k_fs3_eval_f64(Dev, Consts, const double4*, const double*);
__global__ void k_fs3_eval_f32(Dev, Consts, const float4*, const float*);
__global__ void k_fs3_eval_f64_lds(Dev, Consts, const double4*, const double*);
__global__ void k_fs3_eval_f32_lds(Dev, Consts, const float4*, const float*);
__global__ void k_fs3_phi_f64(Dev, Consts, const double4*, const double*, double*);
__global__ void k_fs3_phi_f32(Dev, Consts, const float4*, const float*, double*);
__global__ void k_phi_boxes(Dev, const double*, double*);
int set_orbit_tables(const short* off, const short* mem);
__global__ void k_gram_f64(int, int, const double*, int, int, double*);
__global__ void k_gram_sum(int, int, const double*, double*, int);
template <bool MIXED> __global__ void k_fs2_pair_pass(Dev, Consts);
template <bool MIXED> __global__ void k_fs2_spectator_pass(Dev, Consts);
size_t fs2_lds_bytes(int cap_nb);
size_t fs2_pair_lds_bytes(int cap_nb);
size_t fs3_list_lds_bytes(int cap_nb);
__global__ void k_epilogue(Dev, int, int, int, int, double*, double*, double*, double*, double*, long, long, long,
                           double*, int*);
}  // namespace ccpol3b

using namespace ccpol3b;

namespace {

// Orbit numbering of the FS3 linear coefficients: the label of a site is
// (type, power of the distance opposite that site); a coefficient belongs to
// the multiset of the three labels.  Multisets are numbered 1..364 in order of
// first appearance in the loop site_i -> site_j -> site_k -> p_i -> p_j -> p_k,
// which is the numbering threebody_potential.f90:909-946 produces with fill.
struct OrbitMap {
  std::array<int, 17 * 17 * 17 * 8> idx{};  // [(is*17+js)*17+ks][ip*4+jp*2+kp]
  int count = 0;
  OrbitMap() {
    std::map<std::array<int, 3>, int> seen;
    for (int is = 0; is < 17; ++is)
      for (int js = 0; js < 17; ++js)
        for (int ks = 0; ks < 17; ++ks)
          for (int ip = 0; ip < 2; ++ip)
            for (int jp = 0; jp < 2; ++jp)
              for (int kp = 0; kp < 2; ++kp) {
                std::array<int, 3> key = {site_type(is) * 2 + ip, site_type(js) * 2 + jp,
                                          site_type(ks) * 2 + kp};
                std::sort(key.begin(), key.end());
                auto it = seen.find(key);
                int n;
                if (it == seen.end()) {
                  n = (int)seen.size() + 1;
                  seen.emplace(key, n);
                } else {
                  n = it->second;
                }
                idx[((is * 17 + js) * 17 + ks) * 8 + ip * 4 + jp * 2 + kp] = n;
              }
    count = (int)seen.size();
  }
  int operator()(int is, int js, int ks, int ip, int jp, int kp) const {
    return idx[((is * 17 + js) * 17 + ks) * 8 + ip * 4 + jp * 2 + kp];
  }
};

const OrbitMap& orbit_map() {
  static OrbitMap m;
  return m;
}

int first_site_of_type(int t) {
  for (int s = 0; s < 17; ++s)
    if (site_type(s) == t) return s;
  return -1;
}

// Orbit (0-based) of every entry of the dense table [tx][ty][tz][p][q][s]: all
// sites of one type share their labels, so a representative site per type
// gives the entry's orbit.
const std::vector<int>& dense_orbit() {
  static const std::vector<int> d = [] {
    std::vector<int> v(NCOEF);
    const OrbitMap& om = orbit_map();
    for (int tx = 0; tx < NTYPE; ++tx)
      for (int ty = 0; ty < NTYPE; ++ty)
        for (int tz = 0; tz < NTYPE; ++tz)
          for (int k = 0; k < 8; ++k)
            v[((tx * NTYPE + ty) * NTYPE + tz) * 8 + k] =
                om(first_site_of_type(tx), first_site_of_type(ty), first_site_of_type(tz), k >> 2, (k >> 1) & 1,
                   k & 1) - 1;
    return v;
  }();
  return d;
}

template <typename P>
void dev_free(P*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

struct ccpol3b_ctx {
  int device = 0;
  int precision = CCPOL3B_FP64;
  int s2 = 1, s3 = 1;
  Consts K{};
  std::string err;
  ccpol3b_stats stats{};
  // parameters on the device
  double4* d_pairpar = nullptr;  // [2][289]
  float4* d_pairpar32 = nullptr; // [2][289] FP32 copy for the mixed kernels
  double* d_coef64 = nullptr;    // [2][NCOEF]
  double* d_coef64i = nullptr;   // [2][NCOEF] the same in the interleaved layout of k_fs3_eval_f64
  float* d_coef32s = nullptr;    // [2][NCOEF] FP32, shifted pair-vector basis of the mixed kernels
  double* d_phitrip = nullptr;   // [cap_phi][728] per-triplet coefficient-gradient rows
  long cap_phi = 0;
  ccpol3b_tuning tune{0, 1, 1, 1};
  // creation parameters (Fortran-ordered 17x17 tables, current coefficients),
  // kept so a graph can own a private copy of the context
  std::vector<double> h_tab[6], h_cs[2];
  double alpha3 = 0.0, delta33 = 0.0, rcut3 = 0.0;
  double* d_gpart = nullptr;     // Gram partial sums [row ranges][cols][cols]
  size_t cap_gpart = 0;
  int cu_count = 256;
  // evaluation buffers
  Dev D{};
  int cap_nm = 0, cap_box = 0;
  int* h_totals = nullptr;  // pinned: [0, T_COUNT) mid-call copy, [T_COUNT, 2 T_COUNT) end of call
  int* d_htotals = nullptr; // h_totals + T_COUNT as the device addresses it (null: copied instead)
  // staging for the host entry point
  double* d_in = nullptr;   // x,y,z,xs,ys,zs
  double* d_out = nullptr;  // fxs,fys,fzs (ld=17), fcom2, fcom3, energies
  size_t cap_in = 0, cap_out = 0;
  std::vector<int> h_molbox, h_boxstart;
  std::vector<double> h_boxL;
  // the box description as last copied to the device (empty: none valid); a
  // call whose description is the same copies nothing
  std::vector<int> up_molbox, up_boxstart;
  std::vector<double> up_boxL;
  // cell grids of the uploaded box description (cells a side, first cell)
  std::vector<int> h_cellnc, h_cell0;
  long cap_cells = 0;
  bool cells_zero = false;  // cell_cnt holds zeros (k_cells_build leaves it so)
  hipEvent_t ev[9] = {};
  // the FS3 list build runs on a second stream beside the FS2 passes
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  // the arrays of the last allocation-free call (ccpol3b_eval_device_noalloc),
  // when that was the last call
  Dev nD{};
  bool last_noalloc = false;
  long last_n3 = 0;  // FS3 triplets of the last completed call (sizes the next evaluation grid)
  long version = 0;  // bumped by set_precision, set_coefficients, set_tuning
};

namespace {

int fail(ccpol3b_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                                     \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail(ctx, CCPOL3B_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

#define LAUNCHCHK(ctx, what)                                                                  \
  do {                                                                                        \
    hipError_t e_ = hipGetLastError();                                                        \
    if (e_ != hipSuccess)                                                                     \
      return fail(ctx, CCPOL3B_EHIP, std::string("launch ") + what + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename P>
int dev_alloc(ccpol3b_ctx* c, P*& p, size_t n) {
  dev_free(p);
  if (n == 0) n = 1;
  if (hipMalloc((void**)&p, n * sizeof(P)) != hipSuccess) {
    p = nullptr;
    return fail(c, CCPOL3B_ENOMEM, "device allocation failed");
  }
  return CCPOL3B_OK;
}

void forget_upload(ccpol3b_ctx* c) {
  c->up_molbox.clear();
  c->up_boxstart.clear();
  c->up_boxL.clear();
}

void free_dev(ccpol3b_ctx* c) {
  Dev& D = c->D;
  dev_free(D.cx); dev_free(D.cy); dev_free(D.cz); dev_free(D.site);
  dev_free(D.mol_box); dev_free(D.boxL); dev_free(D.box_start);
  dev_free(D.cell_nc); dev_free(D.cell0); dev_free(D.cell_cnt); dev_free(D.cell_start);
  dev_free(D.cell_mol); dev_free(D.mol_cell);
  D.ncells = 0;
  c->cap_cells = 0;
  dev_free(D.nb3_cnt); dev_free(D.nb3); dev_free(D.up3);
  dev_free(D.nb2_cnt); dev_free(D.nb2); dev_free(D.up2); dev_free(D.nbv2);
  dev_free(D.sl2_cnt); dev_free(D.sl2_off);
  dev_free(D.cnt3s); dev_free(D.off3s); dev_free(D.cnt2s); dev_free(D.off2s);
  dev_free(D.tri); dev_free(D.geo3); dev_free(D.pcnt2); dev_free(D.poff2); dev_free(D.items2);
  dev_free(D.acc); dev_free(D.ew3); dev_free(D.ew2);
  dev_free(D.box_t3); dev_free(D.box_t2); dev_free(D.totals);
  dev_free(D.box_c3); dev_free(D.box_c2); dev_free(D.red3); dev_free(D.red2); dev_free(D.s2pair);
  D.cap_red = 0;
  forget_upload(c);
}

// Per-molecule arrays for nm molecules and nbox boxes.  Sub-list and pair
// arrays are bounded by nm * cap_nb, so they need no mid-call size check.
int ensure_molecules(ccpol3b_ctx* c, int nm, int nbox) {
  Dev& D = c->D;
  int rc;
  if (nm > c->cap_nm) {
    forget_upload(c);
    const int n = std::max(nm, 16);
    const size_t nn = (size_t)n * D.cap_nb;
    if ((rc = dev_alloc(c, D.cx, n)) || (rc = dev_alloc(c, D.cy, n)) || (rc = dev_alloc(c, D.cz, n)) ||
        (rc = dev_alloc(c, D.site, (size_t)n * NS3 * 3)) || (rc = dev_alloc(c, D.mol_box, n)) ||
        (rc = dev_alloc(c, D.cell_mol, n)) || (rc = dev_alloc(c, D.mol_cell, n)) ||
        (rc = dev_alloc(c, D.nb3_cnt, n)) || (rc = dev_alloc(c, D.nb2_cnt, n)) ||
        (rc = dev_alloc(c, D.up3, n)) || (rc = dev_alloc(c, D.up2, n)) ||
        (rc = dev_alloc(c, D.nb3, nn)) || (rc = dev_alloc(c, D.nb2, nn)) || (rc = dev_alloc(c, D.nbv2, 3 * nn)) ||
        (rc = dev_alloc(c, D.sl2_cnt, n)) || (rc = dev_alloc(c, D.sl2_off, n + 1)) ||
        (rc = dev_alloc(c, D.pcnt2, n)) || (rc = dev_alloc(c, D.poff2, n + 1)) ||
        (rc = dev_alloc(c, D.cnt3s, n)) || (rc = dev_alloc(c, D.off3s, n + 1)) ||
        (rc = dev_alloc(c, D.cnt2s, nn)) || (rc = dev_alloc(c, D.off2s, nn + 1)) ||
        (rc = dev_alloc(c, D.s2pair, 2 * nn)) || (rc = dev_alloc(c, D.ew2, (size_t)n * FS2_PARTS_MAX * 2)) ||
        (rc = dev_alloc(c, D.acc, (size_t)n * ACC_W * 2)))
      return rc;
    D.cap_sl2 = D.cap_pairs2 = (long)nn;
    c->cap_nm = n;
  }
  if (nbox > c->cap_box) {
    forget_upload(c);
    if ((rc = dev_alloc(c, D.boxL, nbox)) || (rc = dev_alloc(c, D.box_start, nbox + 1)) ||
        (rc = dev_alloc(c, D.cell_nc, nbox)) || (rc = dev_alloc(c, D.cell0, nbox)) ||
        (rc = dev_alloc(c, D.box_t3, nbox + 1)) || (rc = dev_alloc(c, D.box_t2, nbox + 1)) ||
        (rc = dev_alloc(c, D.box_c3, nbox + 1)) || (rc = dev_alloc(c, D.box_c2, nbox + 1)))
      return rc;
    c->cap_box = nbox;
  }
  if (!D.totals && (rc = dev_alloc(c, D.totals, T_COUNT))) return rc;
  return CCPOL3B_OK;
}

int ensure_lists(ccpol3b_ctx* c, long n3, long ni2) {
  Dev& D = c->D;
  int rc;
  if (n3 > D.cap3) {
    const long n = std::max<long>(n3 + n3 / 8, 1024);
    if ((rc = dev_alloc(c, D.tri, n)) || (rc = dev_alloc(c, D.geo3, n)) || (rc = dev_alloc(c, D.ew3, (size_t)n * 2)))
      return rc;
    D.cap3 = n;
  }
  if (ni2 > D.cap_items2) {
    const long n = std::max<long>(ni2 + ni2 / 8, 1024);
    if ((rc = dev_alloc(c, D.items2, n))) return rc;
    D.cap_items2 = n;
  }
  const long nred = std::max<long>(D.cap3, (long)c->cap_nm * FS2_PARTS_MAX) / RED_CHUNK + D.nbox + 2;
  if (nred > D.cap_red) {
    if ((rc = dev_alloc(c, D.red3, (size_t)nred * 2)) || (rc = dev_alloc(c, D.red2, (size_t)nred * 2))) return rc;
    D.cap_red = nred;
  }
  return CCPOL3B_OK;
}

int waves_grid(long nwaves) { return (int)((nwaves * 64 + 255) / 256); }

// Blocks of a per-call FS3 evaluation (one per triplet): 96 per molecule of
// this call (bulk water has ~61 triplets per molecule, counted by its first
// member), or the last call's count with 1/8 headroom if that is more, at most
// the capacity; so a context whose capacity a large batch has grown does not
// launch millions of empty blocks for a small box.  A list longer than the grid
// is flagged like a capacity overflow (nothing is written) and evaluated again
// with a grid that covers it.  Recorded graphs and the allocation-free entry
// point launch the full capacity.
long fs3_grid(const ccpol3b_ctx* c, int nm, long cap3) {
  const long want = std::max<long>({1024, (long)nm * 96, c->last_n3 + c->last_n3 / 8});
  return std::max<long>(1, std::min(cap3, want));
}

float event_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

// Cell grids of the box description in c->h_*: a periodic box gets nc cells a
// side, nc = floor(L / (r_max (1 + 1e-9))) with r_max the larger of the two
// cutoffs, at most cbrt(8 n_box) (a few molecules per cell or more) and at
// most 1024, when nc reaches CELL_AUTO (tuning neighbour_cells = 1, the
// default) or CELL_MIN (= 2); below that the sweep over the box is cheaper
// than the grid's build and stencil.  Allocates the count arrays.
int setup_cells(ccpol3b_ctx* c, int nbox) {
  Dev& D = c->D;
  const double rmax = std::sqrt(std::max(c->K.rc2_3, c->K.rc2_2));
  c->h_cellnc.assign(nbox, 0);
  c->h_cell0.assign(nbox, 0);
  long ncells = 0;
  for (int b = 0; b < nbox; ++b) {
    const double L = c->h_boxL[b];
    const int nmb = c->h_boxstart[b + 1] - c->h_boxstart[b];
    int nc = 0;
    if (c->tune.neighbour_cells && L > 0.0 && rmax > 0.0 && nmb > 0) {
      const double f = std::min({std::floor(L / (rmax * (1.0 + 1e-9))), std::floor(std::cbrt(8.0 * nmb)), 1024.0});
      if (f >= (c->tune.neighbour_cells == 2 ? CELL_MIN : CELL_AUTO)) nc = (int)f;
    }
    c->h_cellnc[b] = nc;
    c->h_cell0[b] = (int)ncells;
    ncells += (long)nc * nc * nc;
    if (ncells > (1L << 28)) return fail(c, CCPOL3B_ECAPACITY, "too many neighbour-search cells");
  }
  if (ncells > c->cap_cells) {
    int rc;
    if ((rc = dev_alloc(c, D.cell_cnt, ncells)) || (rc = dev_alloc(c, D.cell_start, ncells + 1))) return rc;
    c->cap_cells = ncells;
    c->cells_zero = false;
  }
  D.ncells = (int)ncells;
  return CCPOL3B_OK;
}

// The evaluation proper, on device arrays already in place.  Enqueues the whole
// call on `st` without waiting: grids are sized by the context's capacities
// and every kernel reads the actual list sizes from the device.  A list that
// outgrows its capacity is flagged on the device; the kernels that write the
// caller's outputs then write nothing, and run_checked() grows the capacity and
// evaluates again.  Ends with an asynchronous copy of the counters to the host.
// `fixed`: D is a caller-provided set of device arrays with its capacities and
// box description already in place (the allocation-free entry point); nothing
// is allocated or copied from the host, and c->D is not touched.
int run(ccpol3b_ctx* c, int nm, int nbox, int ld, const double* x, const double* y, const double* z,
        const double* xs, const double* ys, const double* zs, double* fxs, double* fys, double* fzs,
        double* fcom2, double* fcom3, double* energies, hipStream_t st, double* phi = nullptr,
        bool capture = false, Dev* fixed = nullptr) {
  Dev& D = fixed ? *fixed : c->D;
  int rc;
  const int do3 = c->s3, do2 = c->s2;
  // during graph capture nothing may allocate, copy from pageable memory or record timing events
#define EVREC(i) \
  if (!capture) HIPCHK(c, hipEventRecord(c->ev[i], st))
  if (!fixed) {
    if ((rc = ensure_molecules(c, nm, nbox))) return rc;
    D.nm = nm;
    D.nbox = nbox;  // before ensure_lists: the reduction arrays are sized by it
    // triplet capacity: the last call's, or a bulk-water estimate
    if ((rc = ensure_lists(c, std::max<long>(D.cap3, (long)nm * 8), 0))) return rc;
    if (phi && D.cap3 > c->cap_phi) {
      if ((rc = dev_alloc(c, c->d_phitrip, (size_t)D.cap3 * 2 * NLIN3))) return rc;
      c->cap_phi = D.cap3;
    }
  }
  D.fs2_parts = c->tune.fs2_parts;
  D.fs2_pair_parts = c->tune.fs2_pair_parts;
  D.grid3 = capture || fixed ? D.cap3 : fs3_grid(c, nm, D.cap3);
  if (!fixed && !capture &&
      (c->h_molbox != c->up_molbox || c->h_boxL != c->up_boxL || c->h_boxstart != c->up_boxstart)) {
    forget_upload(c);
    HIPCHK(c, hipMemcpyAsync(D.mol_box, c->h_molbox.data(), sizeof(int) * nm, hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(D.boxL, c->h_boxL.data(), sizeof(double) * nbox, hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(D.box_start, c->h_boxstart.data(), sizeof(int) * (nbox + 1),
                             hipMemcpyHostToDevice, st));
    if ((rc = setup_cells(c, nbox))) return rc;
    HIPCHK(c, hipMemcpyAsync(D.cell_nc, c->h_cellnc.data(), sizeof(int) * nbox, hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(D.cell0, c->h_cell0.data(), sizeof(int) * nbox, hipMemcpyHostToDevice, st));
    c->up_molbox = c->h_molbox;
    c->up_boxL = c->h_boxL;
    c->up_boxstart = c->h_boxstart;
  }
  EVREC(0);
  const size_t lds = fs2_lds_bytes(D.cap_nb), lds_p = fs2_pair_lds_bytes(D.cap_nb);
  const size_t lds3 = fs3_list_lds_bytes(D.cap_nb);
  // the neighbour search: FS3 staging, then the cell path's accepted entries
  const size_t ldsn = (do3 ? lds3 : 0) + (D.ncells > 0 ? 2 * sizeof(int) * (size_t)D.cap_nb : 0);
  if (do2 && std::max(lds, lds_p) > 65536)
    return fail(c, CCPOL3B_ECAPACITY, "FS2 neighbour count too large for the FS2 passes");
  if (std::max(do3 ? lds3 : 0, ldsn) > 65536)
    return fail(c, CCPOL3B_ECAPACITY, "neighbour count too large for the neighbour search and the FS3 list build");
  const bool cells = D.ncells > 0 && !fixed;
  if (cells && !c->cells_zero) {
    if (capture) return fail(c, CCPOL3B_EHIP, "cell counts not cleared before graph capture");
    HIPCHK(c, hipMemsetAsync(D.cell_cnt, 0, sizeof(int) * (size_t)D.ncells, st));
    c->cells_zero = true;
  }
  const long nprep = std::max<long>((long)nm * ACC_W, T_COUNT);
  if (cells) c->cells_zero = false;  // until k_cells_build has run
  hipLaunchKernelGGL(k_prep, dim3((nprep + 255) / 256), dim3(256), 0, st, nm, ld, x, y, z, xs, ys, zs, D);
  if (cells) {
    // cell grids: k_prep counted each cell's molecules; their ranges, then the fill
    hipLaunchKernelGGL(k_cells_build<1024>, dim3(1), dim3(1024), 0, st, D);
    LAUNCHCHK(c, "cell grids");
    c->cells_zero = true;
  }
  // neighbour lists with the per-molecule FS3 triplet counts, then every offset
  hipLaunchKernelGGL(k_neighbours, dim3(nm), dim3(256), ldsn, st, D, c->K, do3);
  hipLaunchKernelGGL(k_scan<256>, dim3(1), dim3(256), 0, st, D, 0, do3, do2);
  LAUNCHCHK(c, "neighbour lists");
  EVREC(1);
  // Two branches from here: the FS2 passes on `st`, and the FS3 list fill
  // (triplets in the reference order with their geometry) followed by the FS3
  // evaluation on the context's second stream.  Neither reads what the other
  // writes (both add forces into acc with atomics); `st` waits for the FS3
  // branch before the epilogue.  Under capture the fork and join become graph
  // edges.
#define EVREC_SIDE(i) \
  if (!capture) HIPCHK(c, hipEventRecord(c->ev[i], c->side))
  HIPCHK(c, hipEventRecord(c->fork, st));
  HIPCHK(c, hipStreamWaitEvent(c->side, c->fork, 0));
  if (do2) {
    // pair pass (S_ij, pair-owned forces) before the spectator pass that reads S_ij
    if (c->precision == CCPOL3B_MIXED) {
      hipLaunchKernelGGL(k_fs2_pair_pass<true>, dim3(nm * D.fs2_pair_parts), dim3(256), lds_p, st, D, c->K);
      hipLaunchKernelGGL(k_fs2_spectator_pass<true>, dim3(nm * D.fs2_parts), dim3(256), lds, st, D, c->K);
    } else {
      hipLaunchKernelGGL(k_fs2_pair_pass<false>, dim3(nm * D.fs2_pair_parts), dim3(256), lds_p, st, D, c->K);
      hipLaunchKernelGGL(k_fs2_spectator_pass<false>, dim3(nm * D.fs2_parts), dim3(256), lds, st, D, c->K);
    }
    LAUNCHCHK(c, "FS2");
  }
  EVREC(2);
  hipStream_t s3 = c->side;
  if (do3) hipLaunchKernelGGL(k_fs3_fill, dim3(nm), dim3(256), lds3, s3, D, c->K);
  LAUNCHCHK(c, "FS3 list");
  const long cap3 = D.cap3;
  EVREC_SIDE(8);
  if (do3) {
    const int grid = (int)D.grid3;  // one block per triplet; blocks past the list exit at once
    if (phi) {
      if (c->precision == CCPOL3B_MIXED)
        hipLaunchKernelGGL(k_fs3_phi_f32, dim3(grid), dim3(256), 0, s3, D, c->K, c->d_pairpar32,
                           (const float*)c->d_coef32s, c->d_phitrip);
      else
        hipLaunchKernelGGL(k_fs3_phi_f64, dim3(grid), dim3(256), 0, s3, D, c->K, c->d_pairpar,
                           (const double*)c->d_coef64, c->d_phitrip);
    } else if (c->precision == CCPOL3B_MIXED) {
      if (c->tune.fs3_lds_coef)
        hipLaunchKernelGGL(k_fs3_eval_f32_lds, dim3(grid), dim3(256), 0, s3, D, c->K, c->d_pairpar32,
                           (const float*)c->d_coef32s);
      else
        hipLaunchKernelGGL(k_fs3_eval_f32, dim3(grid), dim3(256), 0, s3, D, c->K, c->d_pairpar32,
                           (const float*)c->d_coef32s);
    } else {
      if (c->tune.fs3_lds_coef)
        hipLaunchKernelGGL(k_fs3_eval_f64_lds, dim3(grid), dim3(256), 0, s3, D, c->K, c->d_pairpar,
                           (const double*)c->d_coef64);
      else
        hipLaunchKernelGGL(k_fs3_eval_f64, dim3(grid), dim3(256), 0, s3, D, c->K, c->d_pairpar,
                           (const double*)c->d_coef64i);
    }
    LAUNCHCHK(c, "k_fs3_eval");
  }
  if (phi) {
    hipLaunchKernelGGL(k_phi_boxes, dim3((2 * NLIN3 + 255) / 256, nbox), dim3(256), 0, s3, D,
                       (const double*)c->d_phitrip, phi);
    LAUNCHCHK(c, "k_phi_boxes");
  }
  EVREC_SIDE(3);
  HIPCHK(c, hipEventRecord(c->join, s3));
  HIPCHK(c, hipStreamWaitEvent(st, c->join, 0));
#undef EVREC_SIDE
  // write-back and the first level of the energy sums (one launch), then the per-box sums
  const long nfin = ((long)nm * ACC_W + 255) / 256;
  const long nch3 = do3 ? cap3 / RED_CHUNK + nbox + 1 : 0;
  const long nch2 = do2 ? (long)nm * D.fs2_parts / RED_CHUNK + nbox + 1 : 0;
  hipLaunchKernelGGL(k_epilogue, dim3(nfin + nch3 + nch2), dim3(256), 0, st, D, ld, do3 ? NS3 : (do2 ? NS2 : 0),
                     do2, do3, fxs, fys, fzs, fcom2, fcom3, nfin, nch3, nch2, energies, c->d_htotals);
  LAUNCHCHK(c, "finish");
  EVREC(4);
  if (!c->d_htotals)
    HIPCHK(c, hipMemcpyAsync(c->h_totals + T_COUNT, D.totals, T_COUNT * sizeof(int), hipMemcpyDeviceToHost, st));
  return CCPOL3B_OK;
#undef EVREC
}

const char* const FXOVER_MSG =
    "a force contribution of 2^30 hartree/bohr or more, or not a number: the geometry is outside the range of the "
    "fixed-point force sums (overlapping sites?)";

// run() until no list outgrew its capacity (at most twice), waiting for the
// stream at the end of each attempt.
int run_checked(ccpol3b_ctx* c, int nm, int nbox, int ld, const double* x, const double* y, const double* z,
                const double* xs, const double* ys, const double* zs, double* fxs, double* fys, double* fzs,
                double* fcom2, double* fcom3, double* energies, hipStream_t st, double* phi = nullptr) {
  Dev& D = c->D;
  c->last_noalloc = false;
  for (int attempt = 0; attempt < 3; ++attempt) {
    int rc = run(c, nm, nbox, ld, x, y, z, xs, ys, zs, fxs, fys, fzs, fcom2, fcom3, energies, st, phi);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(st));
    const int* tot = c->h_totals + T_COUNT;
    if (tot[T_FXOVER]) return fail(c, CCPOL3B_EINVAL, FXOVER_MSG);
    if (tot[T_NBOVER] > 0) {
      // a neighbour list overflowed: grow its capacity (all per-molecule arrays)
      if (tot[T_NBOVER] > 65536) return fail(c, CCPOL3B_ECAPACITY, "neighbour list capacity exceeded");
      D.cap_nb = ((tot[T_NBOVER] + 63) / 64) * 64 + 64;
      c->cap_nm = 0;
      c->cap_box = 0;
      free_dev(c);
      D.cap3 = D.cap_items2 = 0;
      continue;
    }
    if (tot[T_N3] > D.cap3) {
      int rc2 = ensure_lists(c, tot[T_N3], 0);
      if (rc2) return rc2;
      c->last_n3 = tot[T_N3];
      continue;
    }
    if (tot[T_N3] > D.grid3) {  // more triplets than evaluation blocks: again with enough
      c->last_n3 = tot[T_N3];
      continue;
    }
    return CCPOL3B_OK;
  }
  return fail(c, CCPOL3B_ECAPACITY, "list capacity could not be established");
}

void finish_stats(ccpol3b_ctx* c) {
  const int* tot = c->h_totals + T_COUNT;
  c->stats.n_fs3 = c->s3 ? tot[T_N3] : 0;
  c->stats.n_fs2_pairs = c->s2 ? tot[T_NP2] : 0;
  c->stats.n_fs3_sr = tot[T_HIST];
  c->stats.n_fs3_switch = tot[T_HIST + 1];
  c->stats.n_fs3_lr = tot[T_HIST + 2];
  c->stats.n_fs2_items = c->s2 ? tot[T_NI2] : 0;
  c->stats.ms_lists = event_ms(c->ev[0], c->ev[1]);
  c->stats.ms_fs2 = event_ms(c->ev[1], c->ev[2]);
  c->stats.ms_fs3 = event_ms(c->ev[1], c->ev[3]);
  // the epilogue starts when the later of the two branches has ended
  c->stats.ms_reduce = std::min(event_ms(c->ev[2], c->ev[4]), event_ms(c->ev[3], c->ev[4]));
  c->stats.ms_total = event_ms(c->ev[0], c->ev[4]);
  c->stats.ms_fs3_eval = c->stats.n_fs3 > 0 ? event_ms(c->ev[8], c->ev[3]) : 0.0;
  c->stats.n_cells = c->last_noalloc ? 0 : c->D.ncells;
  c->last_n3 = c->s3 ? tot[T_N3] : 0;
}

void set_single_box(ccpol3b_ctx* c, int nm, double box) {
  c->h_molbox.assign(nm, 0);
  c->h_boxstart = {0, nm};
  c->h_boxL = {box > 0.0 ? box / BOHR2A : 0.0};
}

int eval_batch(ccpol3b_ctx* c, int nbox, const int* mol_start, const double* box_len, int ld, const double* x,
               const double* y, const double* z, const double* xs, const double* ys, const double* zs, double* fxs,
               double* fys, double* fzs, double* fcom2, double* fcom3, double* energies, double* phi,
               void* stream);
int check_batch(ccpol3b_ctx* c, int nbox, const int* mol_start, const double* box_len, int ld);
void* noalloc_scratch(int dev);
void set_batch(ccpol3b_ctx* c, int nbox, const int* mol_start, const double* box_len);
int clone_ctx(const ccpol3b_ctx* c, ccpol3b_ctx** out);

// Dense coefficient tables C[X][tx][ty][tz][p][q][s] = cs_X[orbit] (FP64), the
// same entries in the interleaved layout of the FP64 evaluation (cil_index), and
// the FP32 table in the shifted pair-vector basis of the mixed kernels: with v = (e, e r) = T v', v' = (e, e (r - c)), T = [[1, 0], [c, 1]],
// the energy sum_pqs C v_p v_q v_s equals sum C' v'_p v'_q v'_s with
// C'[p'q's'] = sum_pqs C[pqs] T[p][p'] T[q][q'] T[s][s'] (each type block of 8).
int upload_coefficients(ccpol3b_ctx* c, const double* cs_sr, const double* cs_lr) {
  std::vector<double> c64(2 * NCOEF), c64i(2 * NCOEF);
  std::vector<float> c32s(2 * NCOEF);
  const std::vector<int>& dob = dense_orbit();
  for (int X = 0; X < 2; ++X) {
    const double* cs = X ? cs_lr : cs_sr;
    for (int d = 0; d < NCOEF; ++d) {
      c64[(size_t)X * NCOEF + d] = cs[dob[d]];
      const int k = d % 8, tz = d / 8 % NTYPE, tuw = d / (8 * NTYPE);  // d = (tuw NTYPE + tz) 8 + k
      c64i[(size_t)X * NCOEF + cil_index(tz, k) + 2 * tuw] = cs[dob[d]];
    }
    for (int blk = 0; blk < NCOEF / 8; ++blk) {
      double b[8];
      for (int a = 0; a < 8; ++a) b[a] = c64[(size_t)X * NCOEF + blk * 8 + a];
      // index (p q s) = 4 p + 2 q + s; along each axis: new[0] = old[0] + c old[1], new[1] = old[1]
      for (int stride : {4, 2, 1})
        for (int a = 0; a < 8; ++a)
          if ((a & stride) == 0) b[a] += MIXED_VSHIFT * b[a + stride];
      for (int a = 0; a < 8; ++a) c32s[(size_t)X * NCOEF + blk * 8 + a] = (float)b[a];
    }
  }
  if (hipMemcpy(c->d_coef64, c64.data(), sizeof(double) * 2 * NCOEF, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_coef64i, c64i.data(), sizeof(double) * 2 * NCOEF, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_coef32s, c32s.data(), sizeof(float) * 2 * NCOEF, hipMemcpyHostToDevice) != hipSuccess)
    return fail(c, CCPOL3B_EHIP, "coefficient upload");
  return CCPOL3B_OK;
}

}  // namespace

extern "C" {

const char* ccpol3b_status_string(int s) {
  switch (s) {
    case CCPOL3B_OK: return "ok";
    case CCPOL3B_EINVAL: return "invalid argument";
    case CCPOL3B_EHIP: return "HIP runtime error";
    case CCPOL3B_ENOMEM: return "device out of memory";
    case CCPOL3B_ECAPACITY: return "list capacity exceeded";
    case CCPOL3B_ENODEV: return "no gfx950 device";
    default: return "unknown status";
  }
}

const char* ccpol3b_last_error(const ccpol3b_ctx* c) { return c ? c->err.c_str() : "null context"; }

int ccpol3b_orbit_index(int is, int js, int ks, int ip, int jp, int kp) {
  if (is < 1 || is > 17 || js < 1 || js > 17 || ks < 1 || ks > 17 || ip < 0 || ip > 1 || jp < 0 ||
      jp > 1 || kp < 0 || kp > 1)
    return CCPOL3B_EINVAL;
  return orbit_map()(is - 1, js - 1, ks - 1, ip, jp, kp);
}

int ccpol3b_read_data3b(const char* path, double* beta, double* gamma, double* r0, int ld_pair,
                        double* cs, int* nlin) {
  if (!path || !beta || !gamma || !r0 || !cs || !nlin || ld_pair < 17) return CCPOL3B_EINVAL;
  FILE* f = std::fopen(path, "r");
  if (!f) return CCPOL3B_EINVAL;
  char line[512];
  auto next = [&](char* buf) -> bool {
    while (std::fgets(buf, sizeof(line), f)) {
      bool blank = true;
      for (char* q = buf; *q; ++q)
        if (*q != ' ' && *q != '\t' && *q != '\n' && *q != '\r') blank = false;
      if (!blank) return true;
    }
    return false;
  };
  auto fixd = [](char* s) {
    for (char* q = s; *q; ++q)
      if (*q == 'd' || *q == 'D') *q = 'e';
  };
  int rc = CCPOL3B_OK;
  if (!next(line)) rc = CCPOL3B_EINVAL;  // header record
  for (int is = 0; is < 17 && rc == 0; ++is)
    for (int js = 0; js < 17 && rc == 0; ++js) {
      double b, g, r;
      if (!next(line)) { rc = CCPOL3B_EINVAL; break; }
      fixd(line);
      if (std::sscanf(line, "%lf %lf %lf", &b, &g, &r) != 3) { rc = CCPOL3B_EINVAL; break; }
      beta[is + (size_t)ld_pair * js] = b;
      gamma[is + (size_t)ld_pair * js] = g;
      r0[is + (size_t)ld_pair * js] = r;
    }
  int n0 = 0;
  if (rc == 0 && !next(line)) rc = CCPOL3B_EINVAL;  // label record, content ignored
  if (rc == 0 && (!next(line) || std::sscanf(line, "%d", &n0) != 1)) rc = CCPOL3B_EINVAL;
  for (int q = 0; q < n0 && rc == 0; ++q) {
    long idx;
    double v;
    if (!next(line)) { rc = CCPOL3B_EINVAL; break; }
    fixd(line);
    if (std::sscanf(line, "%ld %lf", &idx, &v) != 2) { rc = CCPOL3B_EINVAL; break; }
    if (q < *nlin) cs[q] = v;
  }
  std::fclose(f);
  if (rc == 0) *nlin = n0;
  return rc;
}

int ccpol3b_create(const ccpol3b_params* p, ccpol3b_ctx** out) {
  if (!p || !out) return CCPOL3B_EINVAL;
  *out = nullptr;
  if (p->nlin != CCPOL3B_NLIN || p->ld_pair < 17 || !p->beta_sr || !p->gamma_sr || !p->r0_sr ||
      !p->beta_lr || !p->gamma_lr || !p->r0_lr || !p->cs_sr || !p->cs_lr ||
      (p->precision != CCPOL3B_FP64 && p->precision != CCPOL3B_MIXED))
    return CCPOL3B_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CCPOL3B_ENODEV;
  int dev = p->device;
  if (dev < 0 && hipGetDevice(&dev) != hipSuccess) return CCPOL3B_ENODEV;
  if (dev >= ndev) return CCPOL3B_EINVAL;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return CCPOL3B_ENODEV;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return CCPOL3B_ENODEV;
  if (hipSetDevice(dev) != hipSuccess) return CCPOL3B_ENODEV;

  // the allocation-free entry point's scratch array on this device (its
  // address is resolved once per device; a failure surfaces at that entry point)
  (void)noalloc_scratch(dev);
  ccpol3b_ctx* c = new ccpol3b_ctx();
  c->device = dev;
  c->precision = p->precision;
  c->s2 = p->s2_term ? 1 : 0;
  c->s3 = p->s3_term ? 1 : 0;
  c->cu_count = prop.multiProcessorCount;
  {
    const double* tabs[6] = {p->beta_sr, p->gamma_sr, p->r0_sr, p->beta_lr, p->gamma_lr, p->r0_lr};
    for (int k = 0; k < 6; ++k) {
      c->h_tab[k].resize(NPAIR);
      for (int is = 0; is < NS3; ++is)
        for (int js = 0; js < NS3; ++js) c->h_tab[k][is + NS3 * js] = tabs[k][is + (size_t)p->ld_pair * js];
    }
    c->h_cs[0].assign(p->cs_sr, p->cs_sr + NLIN3);
    c->h_cs[1].assign(p->cs_lr, p->cs_lr + NLIN3);
    c->alpha3 = p->alpha3;
    c->delta33 = p->delta33;
    c->rcut3 = p->rcut3;
  }
  c->D.cap_nb = 256;
  // constants exactly as the reference forms them
  Consts& K = c->K;
  K.alp3 = p->alpha3 * BOHR2A;
  K.delta33 = p->delta33;
  K.idelta = (p->delta33 >= 1.0 && p->delta33 <= 64.0 && p->delta33 == std::floor(p->delta33))
                 ? (int)p->delta33 : 0;
  K.r0dmp3 = 6.20446 / BOHR2A;
  K.r0dmp2 = 7.40446 / BOHR2A;
  K.rc2_3 = std::pow(6.3, 2.0) * std::pow(BOHR2A, -2.0);
  K.rc2_2 = (p->rcut3 * p->rcut3) * std::pow(BOHR2A, -2.0);
  K.rmin = 4.5 / BOHR2A;
  K.rmax = 5.5 / BOHR2A;
  // parameter tables
  std::vector<double4> pairpar(2 * NPAIR);
  std::vector<float4> pairpar32(2 * NPAIR);
  for (int X = 0; X < 2; ++X) {
    const double* b = X ? p->beta_lr : p->beta_sr;
    const double* g = X ? p->gamma_lr : p->gamma_sr;
    const double* r = X ? p->r0_lr : p->r0_sr;
    for (int is = 0; is < 17; ++is)
      for (int js = 0; js < 17; ++js) {
        const size_t q = is + (size_t)p->ld_pair * js;
        pairpar[X * NPAIR + is * 17 + js] = make_double4(b[q], g[q], r[q], 0.0);
        pairpar32[X * NPAIR + is * 17 + js] = make_float4((float)b[q], (float)g[q], (float)r[q], 0.f);
      }
  }
  // orbit -> dense entries (CSR) for the coefficient-gradient kernels
  {
    const std::vector<int>& dob = dense_orbit();
    std::vector<short> off(NLIN3 + 1, 0), mem(NCOEF);
    for (int d = 0; d < NCOEF; ++d) ++off[dob[d] + 1];
    for (int n = 0; n < NLIN3; ++n) off[n + 1] += off[n];
    std::vector<short> fill(off.begin(), off.end() - 1);
    for (int d = 0; d < NCOEF; ++d) mem[fill[dob[d]]++] = (short)d;
    if (set_orbit_tables(off.data(), mem.data()) != 0) {
      delete c;
      return CCPOL3B_EHIP;
    }
  }
  int rc = CCPOL3B_OK;
  if ((rc = dev_alloc(c, c->d_pairpar, 2 * NPAIR)) || (rc = dev_alloc(c, c->d_pairpar32, 2 * NPAIR)) ||
      (rc = dev_alloc(c, c->d_coef64, 2 * NCOEF)) ||
      (rc = dev_alloc(c, c->d_coef64i, 2 * NCOEF)) ||
      (rc = dev_alloc(c, c->d_coef32s, 2 * NCOEF)) ||
      (rc = upload_coefficients(c, p->cs_sr, p->cs_lr))) {
    ccpol3b_destroy(c);
    return rc;
  }
  if (hipMemcpy(c->d_pairpar, pairpar.data(), sizeof(double4) * 2 * NPAIR, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_pairpar32, pairpar32.data(), sizeof(float4) * 2 * NPAIR, hipMemcpyHostToDevice) != hipSuccess ||
      hipHostMalloc((void**)&c->h_totals, 2 * T_COUNT * sizeof(int), hipHostMallocMapped) != hipSuccess) {
    ccpol3b_destroy(c);
    return CCPOL3B_EHIP;
  }
  {
    // the epilogue writes the end-of-call counters straight into the pinned
    // buffer when the device can address it (else they are copied)
    void* dp = nullptr;
    if (hipHostGetDevicePointer(&dp, c->h_totals, 0) == hipSuccess && dp)
      c->d_htotals = static_cast<int*>(dp) + T_COUNT;
    else
      (void)hipGetLastError();
  }
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      ccpol3b_destroy(c);
      return CCPOL3B_EHIP;
    }
  if (hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) {
    ccpol3b_destroy(c);
    return CCPOL3B_EHIP;
  }
  *out = c;
  return CCPOL3B_OK;
}

void ccpol3b_destroy(ccpol3b_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  free_dev(c);
  dev_free(c->d_pairpar);
  dev_free(c->d_pairpar32);
  dev_free(c->d_coef64);
  dev_free(c->d_coef64i);
  dev_free(c->d_coef32s);
  dev_free(c->d_phitrip);
  dev_free(c->d_gpart);
  dev_free(c->d_in);
  dev_free(c->d_out);
  if (c->h_totals) (void)hipHostFree(c->h_totals);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->join) (void)hipEventDestroy(c->join);
  if (c->side) (void)hipStreamDestroy(c->side);
  delete c;
}

int ccpol3b_set_precision(ccpol3b_ctx* c, int precision) {
  if (!c || (precision != CCPOL3B_FP64 && precision != CCPOL3B_MIXED)) return CCPOL3B_EINVAL;
  c->precision = precision;
  ++c->version;
  return CCPOL3B_OK;
}

int ccpol3b_eval_device(ccpol3b_ctx* c, int nm, int ld, const double* x, const double* y, const double* z,
                        const double* xs, const double* ys, const double* zs, double box, double* fxs,
                        double* fys, double* fzs, double* fcom2, double* fcom3, double* energies,
                        void* stream) {
  if (!c || nm < 0 || ld < 17 || (nm > 0 && (!x || !y || !z || !xs || !ys || !zs || !fxs || !fys || !fzs ||
                                             !fcom2 || !fcom3 || !energies)))
    return fail(c, CCPOL3B_EINVAL, "invalid argument");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, CCPOL3B_EHIP, "hipSetDevice");
  hipStream_t st = (hipStream_t)stream;
  set_single_box(c, nm, box);
  if (nm == 0) return CCPOL3B_OK;
  int rc = run_checked(c, nm, 1, ld, x, y, z, xs, ys, zs, fxs, fys, fzs, fcom2, fcom3, energies, st);
  if (rc) return rc;
  finish_stats(c);
  return CCPOL3B_OK;
}

int ccpol3b_gram_device(ccpol3b_ctx* c, int nrow, int ncol, const double* phi, int ld_phi, double* gram,
                        int ld_gram, void* stream) {
  if (!c || nrow < 0 || ncol <= 0 || ld_phi < ncol || ld_gram < ncol || (nrow > 0 && !phi) || !gram)
    return fail(c, CCPOL3B_EINVAL, "invalid argument");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, CCPOL3B_EHIP, "hipSetDevice");
  hipStream_t st = (hipStream_t)stream;
  // row ranges of at least 512 rows (up to 16) so a 728-column matrix fills the chip
  const int nsplit = std::min(16, std::max(1, (nrow + 511) / 512));
  const int rps = ((nrow + nsplit - 1) / nsplit + 15) / 16 * 16;
  const size_t npart = (size_t)nsplit * ncol * ncol;
  int rc;
  if (npart > c->cap_gpart) {
    if ((rc = dev_alloc(c, c->d_gpart, npart))) return rc;
    c->cap_gpart = npart;
  }
  const int nt = (ncol + 63) / 64;
  hipLaunchKernelGGL(k_gram_f64, dim3(nt * (nt + 1) / 2, nsplit), dim3(256), 0, st, nrow, ncol, phi, ld_phi, rps,
                     c->d_gpart);
  const long nn = (long)ncol * ncol;
  hipLaunchKernelGGL(k_gram_sum, dim3((nn + 255) / 256), dim3(256), 0, st, ncol, nsplit,
                     (const double*)c->d_gpart, gram, ld_gram);
  LAUNCHCHK(c, "Gram");
  return CCPOL3B_OK;
}

int ccpol3b_set_coefficients(ccpol3b_ctx* c, const double* cs_sr, const double* cs_lr) {
  if (!c || !cs_sr || !cs_lr) return fail(c, CCPOL3B_EINVAL, "invalid argument");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, CCPOL3B_EHIP, "hipSetDevice");
  // earlier evaluations on any stream still read the old tables
  HIPCHK(c, hipDeviceSynchronize());
  c->h_cs[0].assign(cs_sr, cs_sr + NLIN3);
  c->h_cs[1].assign(cs_lr, cs_lr + NLIN3);
  ++c->version;
  return upload_coefficients(c, cs_sr, cs_lr);
}

int ccpol3b_get_tuning(const ccpol3b_ctx* c, ccpol3b_tuning* t) {
  if (!c || !t) return CCPOL3B_EINVAL;
  *t = c->tune;
  return CCPOL3B_OK;
}

int ccpol3b_set_tuning(ccpol3b_ctx* c, const ccpol3b_tuning* t) {
  if (!c || !t || t->fs2_parts < 1 || t->fs2_parts > FS2_PARTS_MAX || t->fs2_pair_parts < 1 ||
      t->fs2_pair_parts > FS2_PARTS_MAX || t->neighbour_cells < 0 || t->neighbour_cells > 2)
    return fail(c, CCPOL3B_EINVAL, "invalid tuning");
  const int cells = t->neighbour_cells;
  if (cells != c->tune.neighbour_cells) forget_upload(c);  // the cell grids follow the box upload
  c->tune = *t;
  c->tune.fs3_lds_coef = t->fs3_lds_coef ? 1 : 0;
  c->tune.neighbour_cells = cells;
  ++c->version;
  return CCPOL3B_OK;
}

int ccpol3b_autotune(ccpol3b_ctx* c, int nbox, const int* mol_start, const double* box_len, int ld,
                     const double* x, const double* y, const double* z, const double* xs, const double* ys,
                     const double* zs, int reps, void* stream, ccpol3b_tuning* best, double* best_ms) {
  int rc = check_batch(c, nbox, mol_start, box_len, ld);
  if (rc) return rc;
  if (reps < 1 || !x || !y || !z || !xs || !ys || !zs) return fail(c, CCPOL3B_EINVAL, "invalid argument");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, CCPOL3B_EHIP, "hipSetDevice");
  hipStream_t st = (hipStream_t)stream;
  const int nm = mol_start[nbox];
  set_batch(c, nbox, mol_start, box_len);
  double *fs = nullptr, *c2 = nullptr, *c3 = nullptr, *en = nullptr;
  if ((rc = dev_alloc(c, fs, (size_t)3 * ld * nm)) || (rc = dev_alloc(c, c2, (size_t)3 * nm)) ||
      (rc = dev_alloc(c, c3, (size_t)3 * nm)) || (rc = dev_alloc(c, en, (size_t)3 * nbox))) {
    dev_free(fs); dev_free(c2); dev_free(c3); dev_free(en);
    return rc;
  }
  const ccpol3b_tuning keep = c->tune;
  ccpol3b_tuning win = keep;
  double win_ms = 1e300;
  const int parts[] = {1, 2, 4}, pparts[] = {1, 2, 4};
  for (int lds = 0; lds < 2 && rc == CCPOL3B_OK; ++lds)
    for (int pp : parts)
      for (int pw : pparts) {
        if (rc) break;
        c->tune = ccpol3b_tuning{lds, pp, pw, keep.neighbour_cells};
        double tot = 0.0;
        for (int r = 0; r <= reps && rc == CCPOL3B_OK; ++r) {
          rc = run_checked(c, nm, nbox, ld, x, y, z, xs, ys, zs, fs, fs + (size_t)ld * nm,
                           fs + 2 * (size_t)ld * nm, c2, c3, en, st);
          if (rc == CCPOL3B_OK && r > 0) tot += event_ms(c->ev[0], c->ev[4]);
        }
        if (rc == CCPOL3B_OK && tot / reps < win_ms) {
          win_ms = tot / reps;
          win = c->tune;
        }
      }
  dev_free(fs); dev_free(c2); dev_free(c3); dev_free(en);
  if (rc) {
    c->tune = keep;
    return rc;
  }
  c->tune = win;
  ++c->version;
  if (best) *best = win;
  if (best_ms) *best_ms = win_ms;
  return CCPOL3B_OK;
}

int ccpol3b_eval_batch_grad_device(ccpol3b_ctx* c, int nbox, const int* mol_start, const double* box_len, int ld,
                                   const double* x, const double* y, const double* z, const double* xs,
                                   const double* ys, const double* zs, double* fxs, double* fys, double* fzs,
                                   double* fcom2, double* fcom3, double* energies, double* phi, void* stream) {
  if (!phi) return fail(c, CCPOL3B_EINVAL, "invalid argument");
  return eval_batch(c, nbox, mol_start, box_len, ld, x, y, z, xs, ys, zs, fxs, fys, fzs, fcom2, fcom3, energies,
                    phi, stream);
}

int ccpol3b_eval_batch_device(ccpol3b_ctx* c, int nbox, const int* mol_start, const double* box_len, int ld,
                              const double* x, const double* y, const double* z, const double* xs,
                              const double* ys, const double* zs, double* fxs, double* fys, double* fzs,
                              double* fcom2, double* fcom3, double* energies, void* stream) {
  return eval_batch(c, nbox, mol_start, box_len, ld, x, y, z, xs, ys, zs, fxs, fys, fzs, fcom2, fcom3, energies,
                    nullptr, stream);
}
}  // extern "C"

namespace {
int eval_batch(ccpol3b_ctx* c, int nbox, const int* mol_start, const double* box_len, int ld, const double* x,
               const double* y, const double* z, const double* xs, const double* ys, const double* zs, double* fxs,
               double* fys, double* fzs, double* fcom2, double* fcom3, double* energies, double* phi,
               void* stream) {
  int rc0 = check_batch(c, nbox, mol_start, box_len, ld);
  if (rc0) return rc0;
  const int nm = mol_start[nbox];
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, CCPOL3B_EHIP, "hipSetDevice");
  hipStream_t st = (hipStream_t)stream;
  set_batch(c, nbox, mol_start, box_len);
  if (nm == 0) return CCPOL3B_OK;
  int rc = run_checked(c, nm, nbox, ld, x, y, z, xs, ys, zs, fxs, fys, fzs, fcom2, fcom3, energies, st, phi);
  if (rc) return rc;
  finish_stats(c);
  return CCPOL3B_OK;
}

// Validated batch description -> host copies of the box map and lengths.
int check_batch(ccpol3b_ctx* c, int nbox, const int* mol_start, const double* box_len, int ld) {
  if (!c || nbox <= 0 || !mol_start || !box_len || ld < 17) return fail(c, CCPOL3B_EINVAL, "invalid argument");
  if (mol_start[0] != 0 || mol_start[nbox] < 0) return fail(c, CCPOL3B_EINVAL, "mol_start must start at 0");
  for (int b = 0; b < nbox; ++b)
    if (mol_start[b + 1] < mol_start[b]) return fail(c, CCPOL3B_EINVAL, "mol_start not ascending");
  return CCPOL3B_OK;
}

void set_batch(ccpol3b_ctx* c, int nbox, const int* mol_start, const double* box_len) {
  const int nm = mol_start[nbox];
  c->h_molbox.resize(nm);
  c->h_boxstart.assign(mol_start, mol_start + nbox + 1);
  c->h_boxL.resize(nbox);
  for (int b = 0; b < nbox; ++b) {
    c->h_boxL[b] = box_len[b] > 0.0 ? box_len[b] / BOHR2A : 0.0;
    for (int m = mol_start[b]; m < mol_start[b + 1]; ++m) c->h_molbox[m] = b;
  }
}

// A new context with the same parameters, precision, switches and tuning.
int clone_ctx(const ccpol3b_ctx* c, ccpol3b_ctx** out) {
  ccpol3b_params p{};
  p.beta_sr = c->h_tab[0].data();
  p.gamma_sr = c->h_tab[1].data();
  p.r0_sr = c->h_tab[2].data();
  p.beta_lr = c->h_tab[3].data();
  p.gamma_lr = c->h_tab[4].data();
  p.r0_lr = c->h_tab[5].data();
  p.ld_pair = 17;
  p.cs_sr = c->h_cs[0].data();
  p.cs_lr = c->h_cs[1].data();
  p.nlin = CCPOL3B_NLIN;
  p.alpha3 = c->alpha3;
  p.delta33 = c->delta33;
  p.rcut3 = c->rcut3;
  p.s2_term = c->s2;
  p.s3_term = c->s3;
  p.precision = c->precision;
  p.device = c->device;
  int rc = ccpol3b_create(&p, out);
  if (rc == CCPOL3B_OK) (*out)->tune = c->tune;
  return rc;
}
}  // namespace

extern "C" {

int ccpol3b_threebody_term(ccpol3b_ctx* c, int nm, int ld, const double* x, const double* y, const double* z,
                           const double* xs, const double* ys, const double* zs, double box, double* qfs2,
                           double* qfs3, double* fxs, double* fys, double* fzs, double* fxfs2, double* fyfs2,
                           double* fzfs2, double* fxfs3, double* fyfs3, double* fzfs3, double* txfs2,
                           double* tyfs2, double* tzfs2, double* txfs3, double* tyfs3, double* tzfs3,
                           double* vir) {
  if (!c || nm < 0 || ld < 17 || !qfs2 || !qfs3 || !vir ||
      (nm > 0 && (!x || !y || !z || !xs || !ys || !zs || !fxs || !fys || !fzs || !fxfs2 || !fyfs2 ||
                  !fzfs2 || !fxfs3 || !fyfs3 || !fzfs3)))
    return fail(c, CCPOL3B_EINVAL, "invalid argument");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, CCPOL3B_EHIP, "hipSetDevice");
  if (nm == 0) return CCPOL3B_OK;
  const int ldd = NS3;  // compact device layout of site arrays
  const size_t nin = 3 * (size_t)nm + 3 * (size_t)ldd * nm;
  const size_t nout = 3 * (size_t)ldd * nm + 6 * (size_t)nm + 3;
  int rc;
  if (nin > c->cap_in) {
    if ((rc = dev_alloc(c, c->d_in, nin))) return rc;
    c->cap_in = nin;
  }
  if (nout > c->cap_out) {
    if ((rc = dev_alloc(c, c->d_out, nout))) return rc;
    c->cap_out = nout;
  }
  // pack the inputs: COM and the first 17 rows of the site arrays
  std::vector<double> hin(nin);
  std::memcpy(hin.data(), x, sizeof(double) * nm);
  std::memcpy(hin.data() + nm, y, sizeof(double) * nm);
  std::memcpy(hin.data() + 2 * nm, z, sizeof(double) * nm);
  const double* srcs[3] = {xs, ys, zs};
  for (int a = 0; a < 3; ++a)
    for (int m = 0; m < nm; ++m)
      std::memcpy(hin.data() + 3 * nm + (size_t)a * ldd * nm + (size_t)ldd * m, srcs[a] + (size_t)ld * m,
                  sizeof(double) * ldd);
  hipStream_t st = nullptr;
  HIPCHK(c, hipEventRecord(c->ev[5], st));
  HIPCHK(c, hipMemcpyAsync(c->d_in, hin.data(), sizeof(double) * nin, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemsetAsync(c->d_out, 0, sizeof(double) * nout, st));
  double* din = c->d_in;
  double* dout = c->d_out;
  set_single_box(c, nm, box);
  rc = run_checked(c, nm, 1, ldd, din, din + nm, din + 2 * nm, din + 3 * nm, din + 3 * nm + (size_t)ldd * nm,
           din + 3 * nm + 2 * (size_t)ldd * nm, dout, dout + (size_t)ldd * nm, dout + 2 * (size_t)ldd * nm,
           dout + 3 * (size_t)ldd * nm, dout + 3 * (size_t)ldd * nm + 3 * nm, dout + 3 * (size_t)ldd * nm + 6 * nm,
           st);
  if (rc) return rc;
  std::vector<double> hout(nout);
  HIPCHK(c, hipEventRecord(c->ev[6], st));
  HIPCHK(c, hipMemcpyAsync(hout.data(), dout, sizeof(double) * nout, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipEventRecord(c->ev[7], st));
  HIPCHK(c, hipStreamSynchronize(st));
  finish_stats(c);
  c->stats.ms_h2d = event_ms(c->ev[5], c->ev[0]);
  c->stats.ms_d2h = event_ms(c->ev[6], c->ev[7]);
  // scatter back with the reference's output semantics
  const double* f[3] = {hout.data(), hout.data() + (size_t)ldd * nm, hout.data() + 2 * (size_t)ldd * nm};
  double* dst[3] = {fxs, fys, fzs};
  const int nsite_out = c->s3 ? NS3 : (c->s2 ? NS2 : 0);
  for (int a = 0; a < 3; ++a)
    for (int m = 0; m < nm; ++m)
      for (int s = 0; s < nsite_out; ++s) dst[a][(size_t)ld * m + s] += f[a][(size_t)ldd * m + s];
  const double* c2 = hout.data() + 3 * (size_t)ldd * nm;
  const double* c3 = c2 + 3 * nm;
  const double* en = c3 + 3 * nm;
  if (c->s2) {
    for (int m = 0; m < nm; ++m) {
      fxfs2[m] = c2[m];
      fyfs2[m] = c2[nm + m];
      fzfs2[m] = c2[2 * nm + m];
      if (txfs2) txfs2[m] = 0.0;
      if (tyfs2) tyfs2[m] = 0.0;
      if (tzfs2) tzfs2[m] = 0.0;
    }
    *qfs2 = en[0];
  }
  if (c->s3) {
    for (int m = 0; m < nm; ++m) {
      fxfs3[m] = c3[m];
      fyfs3[m] = c3[nm + m];
      fzfs3[m] = c3[2 * nm + m];
      if (txfs3) txfs3[m] = 0.0;
      if (tyfs3) tyfs3[m] = 0.0;
      if (tzfs3) tzfs3[m] = 0.0;
    }
    *qfs3 = en[1];
  }
  *vir += en[2];
  return CCPOL3B_OK;
}

int ccpol3b_last_stats(const ccpol3b_ctx* c, ccpol3b_stats* s) {
  if (!c || !s) return CCPOL3B_EINVAL;
  *s = c->stats;
  return CCPOL3B_OK;
}

long ccpol3b_get_fs3_list(ccpol3b_ctx* c, int* ijk, long cap) {
  if (!c || (cap > 0 && !ijk)) return CCPOL3B_EINVAL;
  const long n = c->stats.n_fs3;
  const long m = std::min(n, cap);
  if (m > 0) {
    std::vector<int4> tmp(m);
    const int4* tri = c->last_noalloc ? c->nD.tri : c->D.tri;
    if (hipMemcpy(tmp.data(), tri, sizeof(int4) * m, hipMemcpyDeviceToHost) != hipSuccess)
      return fail(c, CCPOL3B_EHIP, "copy of the FS3 list");
    for (long t = 0; t < m; ++t) {
      ijk[3 * t] = tmp[t].x;
      ijk[3 * t + 1] = tmp[t].y;
      ijk[3 * t + 2] = tmp[t].z;
    }
  }
  return n;
}

long ccpol3b_get_fs2_list(ccpol3b_ctx* c, int* kij, long cap) {
  if (!c || (cap > 0 && !kij)) return CCPOL3B_EINVAL;
  if (c->last_noalloc)
    return fail(c, CCPOL3B_EINVAL, "the FS2 list is exported after ccpol3b_eval_device, not after the "
                                   "allocation-free call");
  if (!c->s2 || c->D.nm == 0) return 0;
  // The evaluation does not materialise the FS2 items; build them from the
  // neighbour lists of the last call, in the reference's order.
  Dev& D = c->D;
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, CCPOL3B_EHIP, "hipSetDevice");
  const hipStream_t st = nullptr;
  const long nsl2 = c->h_totals[T_COUNT + T_SL2];
  if (nsl2 > 0) hipLaunchKernelGGL(k_fs2_items, dim3(waves_grid(nsl2)), dim3(256), 0, st, D, c->K, 0);
  hipLaunchKernelGGL(k_scan<1024>, dim3(1), dim3(1024), 0, st, D, 2, 0, 1);
  LAUNCHCHK(c, "FS2 list count");
  int ni = 0;
  HIPCHK(c, hipMemcpy(&ni, D.totals + T_NI2L, sizeof(int), hipMemcpyDeviceToHost));
  int rc;
  if ((rc = ensure_lists(c, 0, ni))) return rc;
  if (nsl2 > 0) hipLaunchKernelGGL(k_fs2_items, dim3(waves_grid(nsl2)), dim3(256), 0, st, D, c->K, 1);
  LAUNCHCHK(c, "FS2 list fill");
  const long m = std::min<long>(ni, cap);
  if (m > 0) {
    std::vector<int4> tmp(m);
    HIPCHK(c, hipMemcpy(tmp.data(), D.items2, sizeof(int4) * m, hipMemcpyDeviceToHost));
    for (long t = 0; t < m; ++t) {
      kij[3 * t] = tmp[t].x;
      kij[3 * t + 1] = tmp[t].y;
      kij[3 * t + 2] = tmp[t].z;
    }
  } else {
    HIPCHK(c, hipDeviceSynchronize());
  }
  return ni;
}

struct ccpol3b_graph {
  ccpol3b_ctx* c = nullptr;  // private context: the recorded kernels hold its device pointers
  hipStream_t st = nullptr;  // capture stream
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
};

void ccpol3b_graph_destroy(ccpol3b_graph* g) {
  if (!g) return;
  if (g->c) (void)hipSetDevice(g->c->device);
  if (g->x) (void)hipGraphExecDestroy(g->x);
  if (g->g) (void)hipGraphDestroy(g->g);
  if (g->st) (void)hipStreamDestroy(g->st);
  if (g->c) ccpol3b_destroy(g->c);
  delete g;
}

int ccpol3b_graph_create(ccpol3b_ctx* c, int nbox, const int* mol_start, const double* box_len, int ld,
                         const double* x, const double* y, const double* z, const double* xs, const double* ys,
                         const double* zs, double* fxs, double* fys, double* fzs, double* fcom2, double* fcom3,
                         double* energies, double* phi, double headroom, void* stream, ccpol3b_graph** out) {
  if (!out) return fail(c, CCPOL3B_EINVAL, "invalid argument");
  *out = nullptr;
  int rc = check_batch(c, nbox, mol_start, box_len, ld);
  if (rc) return rc;
  const int nm = mol_start[nbox];
  if (nm <= 0 || !x || !y || !z || !xs || !ys || !zs || !fxs || !fys || !fzs || !fcom2 || !fcom3 || !energies ||
      !(headroom >= 0.0 && headroom <= 16.0))
    return fail(c, CCPOL3B_EINVAL, "invalid argument");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, CCPOL3B_EHIP, "hipSetDevice");
  ccpol3b_graph* g = new ccpol3b_graph();
  if ((rc = clone_ctx(c, &g->c))) {
    delete g;
    return fail(c, rc, "graph context");
  }
  ccpol3b_ctx* gc = g->c;
  hipStream_t st = (hipStream_t)stream;
  // one checked evaluation sizes every list; its outputs are valid
  set_batch(gc, nbox, mol_start, box_len);
  rc = run_checked(gc, nm, nbox, ld, x, y, z, xs, ys, zs, fxs, fys, fzs, fcom2, fcom3, energies, st, phi);
  if (rc == CCPOL3B_OK) {
    finish_stats(gc);
    c->stats = gc->stats;
    const long want = (long)std::ceil((double)std::max<long>(gc->D.cap3, gc->stats.n_fs3) * (1.0 + headroom));
    rc = ensure_lists(gc, want, 0);
    if (rc == CCPOL3B_OK && phi && gc->D.cap3 > gc->cap_phi) {
      rc = dev_alloc(gc, gc->d_phitrip, (size_t)gc->D.cap3 * 2 * NLIN3);
      if (rc == CCPOL3B_OK) gc->cap_phi = gc->D.cap3;
    }
  }
  if (rc == CCPOL3B_OK && hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking) != hipSuccess) rc = CCPOL3B_EHIP;
  if (rc == CCPOL3B_OK && hipStreamSynchronize(st) != hipSuccess) rc = CCPOL3B_EHIP;
  if (rc == CCPOL3B_OK) {
    if (hipStreamBeginCapture(g->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      rc = CCPOL3B_EHIP;
    } else {
      rc = run(gc, nm, nbox, ld, x, y, z, xs, ys, zs, fxs, fys, fzs, fcom2, fcom3, energies, g->st, phi, true);
      hipGraph_t gr = nullptr;
      if (hipStreamEndCapture(g->st, &gr) != hipSuccess && rc == CCPOL3B_OK) rc = CCPOL3B_EHIP;
      g->g = gr;
      if (rc == CCPOL3B_OK && (!gr || hipGraphInstantiate(&g->x, gr, nullptr, nullptr, 0) != hipSuccess))
        rc = CCPOL3B_EHIP;
    }
  }
  if (rc) {
    std::string msg = gc->err.empty() ? "graph capture failed" : gc->err;
    ccpol3b_graph_destroy(g);
    return fail(c, rc, msg);
  }
  *out = g;
  return CCPOL3B_OK;
}

int ccpol3b_graph_launch(ccpol3b_graph* g, void* stream) {
  if (!g || !g->x) return CCPOL3B_EINVAL;
  if (hipSetDevice(g->c->device) != hipSuccess) return CCPOL3B_EHIP;
  if (hipGraphLaunch(g->x, (hipStream_t)stream) != hipSuccess) return fail(g->c, CCPOL3B_EHIP, "hipGraphLaunch");
  return CCPOL3B_OK;
}

}  // extern "C"

long ccpol3b::context_version(const ccpol3b_ctx* c) { return c ? c->version : -1; }

const int* ccpol3b::graph_device_totals(const ccpol3b_graph* g, long* cap3) {
  if (!g || !g->c) return nullptr;
  if (cap3) *cap3 = g->c->D.cap3;
  return g->c->D.totals;
}

extern "C" {

int ccpol3b_graph_status(const ccpol3b_graph* g) {
  if (!g || !g->c) return CCPOL3B_EINVAL;
  const int* tot = g->c->h_totals + T_COUNT;
  if (tot[T_FXOVER]) return CCPOL3B_EINVAL;
  if (tot[T_NBOVER] > 0 || tot[T_N3] > g->c->D.cap3) return CCPOL3B_ECAPACITY;
  return CCPOL3B_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free entry point
// ccpol3b_eval_device_noalloc evaluates one box with every per-call device
// array carved from `scratch`, a file-scope device array of fixed size, instead
// of the context's allocations: no hipMalloc/hipFree, no host-to-device copy.
// Each GPU holds its own copy of the array; its address is resolved once per
// device (hipGetSymbolAddress with that device current) and kept in
// noalloc_ptr[].  Host code uses that address, never the array's name.  All
// calls on one GPU share the array (it keeps the last call's contents; k_prep
// clears the accumulators and counters each call), so calls that overlap on
// one GPU race.
namespace {

constexpr int NA_NM_MAX = 1024;            // molecules of one call
constexpr int NA_CAP_NB = 128;             // FS3 / FS2 neighbours per molecule
constexpr long NA_CAP3 = 96L * NA_NM_MAX;  // FS3 triplets
constexpr size_t NA_SCRATCH_BYTES = size_t(32) << 20;
constexpr int NA_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

std::mutex noalloc_mutex;
void* noalloc_ptr[NA_MAX_DEVICES] = {};

// Lays the arrays of one evaluation (capacities NA_*) out from `base`, each
// aligned to 256 bytes, and returns the bytes used; with base == nullptr it
// only measures.
size_t noalloc_carve(unsigned char* base, Dev& D) {
  size_t off = 0;
  auto take = [&](auto*& p, size_t n) {
    using P = std::remove_reference_t<decltype(*p)>;
    p = base ? reinterpret_cast<P*>(base + off) : nullptr;
    off += (std::max<size_t>(n, 1) * sizeof(P) + 255) / 256 * 256;
  };
  const size_t n = NA_NM_MAX, nn = n * NA_CAP_NB, nbox = 1;
  const long nred = std::max<long>(NA_CAP3, (long)NA_NM_MAX * FS2_PARTS_MAX) / RED_CHUNK + (long)nbox + 2;
  D.cap_nb = NA_CAP_NB;
  D.cap3 = NA_CAP3;
  D.cap_pairs2 = D.cap_sl2 = (long)nn;
  D.cap_items2 = 0;
  D.cap_red = nred;
  D.items2 = nullptr;
  take(D.cx, n); take(D.cy, n); take(D.cz, n); take(D.site, n * NS3 * 3);
  take(D.mol_box, n); take(D.boxL, nbox); take(D.box_start, nbox + 1);
  take(D.cell_nc, nbox);  // 0: the sweep (no cell grid)
  take(D.nb3_cnt, n); take(D.nb3, nn); take(D.up3, n);
  take(D.nb2_cnt, n); take(D.nb2, nn); take(D.up2, n); take(D.nbv2, 3 * nn);
  take(D.sl2_cnt, n); take(D.sl2_off, n + 1);
  take(D.cnt3s, n); take(D.off3s, n + 1);
  take(D.cnt2s, nn); take(D.off2s, nn + 1);
  take(D.tri, (size_t)NA_CAP3); take(D.geo3, (size_t)NA_CAP3);
  take(D.pcnt2, n); take(D.poff2, n + 1);
  take(D.acc, n * ACC_W * 2); take(D.ew3, 2 * (size_t)NA_CAP3); take(D.ew2, n * FS2_PARTS_MAX * 2);
  take(D.box_t3, nbox + 1); take(D.box_t2, nbox + 1); take(D.box_c3, nbox + 1); take(D.box_c2, nbox + 1);
  take(D.red3, 2 * (size_t)nred); take(D.red2, 2 * (size_t)nred);
  take(D.s2pair, 2 * nn);
  take(D.totals, T_COUNT);
  return off;
}

// The box description of a single box: every molecule in box 0 of length L (bohr).
__global__ void k_single_box(Dev D, int nm, double L) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nm) D.mol_box[t] = 0;
  if (t == 0) {
    D.boxL[0] = L;
    D.box_start[0] = 0;
    D.box_start[1] = nm;
    D.cell_nc[0] = 0;
  }
}

// Address of this device's copy of `scratch` (dev must be the current device);
// resolved on first use and kept.  nullptr if it cannot be resolved.
void* noalloc_scratch(int dev) {
  if (dev < 0 || dev >= NA_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(noalloc_mutex);
  if (!noalloc_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess && p) noalloc_ptr[dev] = p;
  }
  return noalloc_ptr[dev];
}

}  // namespace

extern "C" int ccpol3b_eval_device_noalloc(ccpol3b_ctx* c, int nm, int ld, const double* x, const double* y,
                                           const double* z, const double* xs, const double* ys,
                                           const double* zs, double box, double* fxs, double* fys, double* fzs,
                                           double* fcom2, double* fcom3, double* energies, void* stream) {
  if (!c || nm < 0 || ld < 17 || (nm > 0 && (!x || !y || !z || !xs || !ys || !zs || !fxs || !fys || !fzs ||
                                             !fcom2 || !fcom3 || !energies)))
    return fail(c, CCPOL3B_EINVAL, "invalid argument");
  if (nm > NA_NM_MAX)
    return fail(c, CCPOL3B_ECAPACITY, "the allocation-free entry point takes at most 1024 molecules");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, CCPOL3B_EHIP, "hipSetDevice");
  unsigned char* base = static_cast<unsigned char*>(noalloc_scratch(c->device));
  if (!base) return fail(c, CCPOL3B_EHIP, "address of the allocation-free scratch array");
  Dev D{};
  if (noalloc_carve(nullptr, D) > NA_SCRATCH_BYTES)
    return fail(c, CCPOL3B_EHIP, "allocation-free scratch layout exceeds its array");
  if (nm == 0) return CCPOL3B_OK;
  noalloc_carve(base, D);
  D.nm = nm;
  D.nbox = 1;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_single_box, dim3((nm + 255) / 256), dim3(256), 0, st, D, nm,
                     box > 0.0 ? box / BOHR2A : 0.0);
  LAUNCHCHK(c, "box description");
  int rc = run(c, nm, 1, ld, x, y, z, xs, ys, zs, fxs, fys, fzs, fcom2, fcom3, energies, st, nullptr, false, &D);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(st));
  const int* tot = c->h_totals + T_COUNT;
  if (tot[T_FXOVER]) return fail(c, CCPOL3B_EINVAL, FXOVER_MSG);
  if (tot[T_NBOVER] > 0)
    return fail(c, CCPOL3B_ECAPACITY, "a molecule has more than 128 neighbours (allocation-free entry point)");
  if (tot[T_N3] > D.cap3)
    return fail(c, CCPOL3B_ECAPACITY, "more than 98304 FS3 triplets (allocation-free entry point)");
  c->nD = D;
  c->last_noalloc = true;
  finish_stats(c);
  return CCPOL3B_OK;
}
