// Rigid-water NVE driver on the device, for the three-body forces of this
// library (the caller's two-body, Ewald and induction terms are not part of it).
//
// Scheme.  Centres of mass: velocity Verlet.  Orientations: the symplectic
// NO_SQUISH splitting of Miller, Eleftheriou, Pattnaik, Ndirango, Newns and
// Martyna (J. Chem. Phys. 116, 8649 (2002)), the integrator the reference's
// nve_1 uses (evolve.f90:621-1062, 8107-8204).  In the paper's notation the
// orientation is a unit quaternion q with conjugate momentum p, related to the
// body-frame angular momentum j by p = 2 sum_k j_k P_k q and j_k = (P_k q).p / 2,
// where P_1..P_3 are the signed permutations of the quaternion components
// below.  The free-rotor flow splits into exact rotations about the principal
// axes: rotating about axis k for a time t turns (q, p) in the planes of
// (q, P_k q) and (p, P_k p) by the angle t (p . P_k q) / (4 I_k).  One
// sub-step applies axes 3, 2, 1, 2, 3 for t/2, t/2, t, t/2, t/2; a time step
// is NSUB such sub-steps.  Torques act on j directly: the half kick
// p += dt/2 * 2 sum_k tau_k P_k q at fixed q is j += dt/2 * tau (body frame).
//
// Per step, two launches:
//   three-body forces:  the recorded evaluation (ccpol3b_graph_launch);
//   k_md_edge:          second half of the step (half kick, COM wrapped into
//                       the box, kinetic energies) and its record (E3, Kt, Kr
//                       and their sum), then the first half of the next step
//                       (half kick of v and j, drift of x, NO_SQUISH rotation)
//                       with the lab-frame site offsets R(q) c_s and zeroed
//                       site forces for the next evaluation.
// State per molecule: COM (A), quaternion, COM velocity (A/ps) and body-frame
// angular momentum (amu A^2/ps); forces in 10 J mol^-1 A^-1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/ccpol3b.h"
#include "ccpol3b_internal.h"

namespace ccpol3b {

constexpr int NSUB = 10;  // NO_SQUISH sub-steps per time step

// (P_k q)[a] = sign * q[src]: the three signed permutations of the paper, as
// (source component, sign) per output component.
struct SignedPerm {
  int src[4];
  double sign[4];
};
constexpr SignedPerm kPerm[3] = {
    {{1, 0, 3, 2}, {-1.0, 1.0, 1.0, -1.0}},   // P_1
    {{2, 3, 0, 1}, {-1.0, -1.0, 1.0, 1.0}},   // P_2
    {{3, 2, 1, 0}, {-1.0, 1.0, -1.0, 1.0}},   // P_3
};

// The symmetric sequence of one sub-step: (axis, fraction of the sub-step).
struct RotorStep {
  int axis;
  double frac;
};
constexpr RotorStep kRotorSeq[5] = {{2, 0.5}, {1, 0.5}, {0, 1.0}, {1, 0.5}, {2, 0.5}};

__device__ __forceinline__ void permute(int k, const double* v, double* out) {
#pragma unroll
  for (int a = 0; a < 4; ++a) out[a] = kPerm[k].sign[a] * v[kPerm[k].src[a]];
}

__device__ __forceinline__ double dot4(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

// Exact free rotation about principal axis k for time t.
__device__ __forceinline__ void free_rotation(int k, double t, double inertia, double* q, double* p) {
  double pq[4], pp[4];
  permute(k, q, pq);
  permute(k, p, pp);
  const double angle = t * dot4(p, pq) / (4.0 * inertia);
  double s, c;
  sincos(angle, &s, &c);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    q[a] = c * q[a] + s * pq[a];
    p[a] = c * p[a] + s * pp[a];
  }
}

// NO_SQUISH propagation of (q, j) over dt with principal moments pmi.
__device__ void rotate_rigid_body(double dt, const double* pmi, double* q, double* j) {
  double p[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double pq[4];
    permute(k, q, pq);
#pragma unroll
    for (int a = 0; a < 4; ++a) p[a] += 2.0 * j[k] * pq[a];
  }
  const double t = dt / NSUB;
  for (int n = 0; n < NSUB; ++n) {
#pragma unroll
    for (int r = 0; r < 5; ++r) free_rotation(kRotorSeq[r].axis, kRotorSeq[r].frac * t, pmi[kRotorSeq[r].axis], q, p);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double pq[4];
    permute(k, q, pq);
    j[k] = 0.5 * dot4(p, pq);
  }
}

// Rotation matrix of a unit quaternion (body -> lab), row major; the same
// matrix as quat_to_rot (molecular_sites.f90:95-111).
__device__ __forceinline__ void rotation_matrix(const double* q, double* R) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  R[0] = w * w + x * x - y * y - z * z;
  R[1] = 2.0 * (x * y - w * z);
  R[2] = 2.0 * (x * z + w * y);
  R[3] = 2.0 * (x * y + w * z);
  R[4] = w * w - x * x + y * y - z * z;
  R[5] = 2.0 * (y * z - w * x);
  R[6] = 2.0 * (x * z - w * y);
  R[7] = 2.0 * (y * z + w * x);
  R[8] = w * w - x * x - y * y + z * z;
}

struct MdDev {
  int nm, ns, ld;
  double dt, mass, pmi[3], box;
  double* com;    // [3][nm]
  double* q;      // [4][nm]
  double* vel;    // [3][nm]
  double* jb;     // [3][nm]
  double* pc;     // [ns][3] principal-frame sites
  double* site;   // [3][nm][ld] lab-frame offsets
  double* fsite;  // [3][nm][ld]
  double* fc2;    // [3][nm]
  double* fc3;    // [3][nm]
  double* en;     // [3] qfs2, qfs3, vir increment
  double* ke;     // [2][nm] translational, rotational kinetic energy per molecule
  double* rec;    // [nsteps][4] E3, Kt, Kr, E3 + Kt + Kr (kcal/mol)
  int* failed;    // [0] set when a force evaluation wrote nothing (FAIL_LISTS, FAIL_FORCE);
                  // [1] blocks of k_md_edge done (the last one writes the step record)
  const int* ftot;  // the force evaluation's counters (see graph_device_totals)
  long fcap3;
};

// One rigid body in registers.
struct Body {
  double x[3], v[3], q[4], j[3];
};

__device__ __forceinline__ Body load_body(const MdDev& M, int m) {
  Body b;
  const int nm = M.nm;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    b.x[c] = M.com[c * nm + m];
    b.v[c] = M.vel[c * nm + m];
    b.j[c] = M.jb[c * nm + m];
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) b.q[c] = M.q[c * nm + m];
  return b;
}

__device__ __forceinline__ void store_body(const MdDev& M, int m, const Body& b) {
  const int nm = M.nm;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    M.com[c * nm + m] = b.x[c];
    M.vel[c * nm + m] = b.v[c];
    M.jb[c * nm + m] = b.j[c];
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) M.q[c * nm + m] = b.q[c];
}

// Why the last force evaluation wrote nothing: 0 it did, FAIL_LISTS a list
// outgrew its capacity, FAIL_FORCE a force left the fixed-point range.
enum { FAIL_LISTS = 1, FAIL_FORCE = 2 };
__device__ __forceinline__ int evaluation_failure(const MdDev& M) {
  if (M.ftot[T_FXOVER]) return FAIL_FORCE;
  return (M.ftot[T_NBOVER] > 0 || M.ftot[T_N3] > M.fcap3) ? FAIL_LISTS : 0;
}

// Half kick of v and j with the force F and body-frame torque tb.
__device__ __forceinline__ void half_kick(const MdDev& M, Body& b, const double* F, const double* tb) {
  const double h = 0.5 * M.dt;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    b.v[c] += h * F[c] / M.mass;
    b.j[c] += h * tb[c];
  }
}

// Lab-frame site offsets R(q) c_s of molecule m (zero in the padding rows of
// the ld-row layout) and zeroed site forces.
__device__ __forceinline__ void write_sites(const MdDev& M, int m, const double* q) {
  const long plane = (long)M.nm * M.ld, base = (long)m * M.ld;
  double R[9];
  rotation_matrix(q, R);
  for (int st = 0; st < M.ld; ++st) {
    double lab[3] = {0.0, 0.0, 0.0};
    if (st < M.ns) {
      const double* c = M.pc + 3 * st;
#pragma unroll
      for (int a = 0; a < 3; ++a) lab[a] = R[3 * a] * c[0] + R[3 * a + 1] * c[1] + R[3 * a + 2] * c[2];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      M.site[a * plane + base + st] = lab[a];
      M.fsite[a * plane + base + st] = 0.0;
    }
  }
}

// Sites of every molecule (before the first evaluation of a run).
__global__ void k_md_sites(MdDev M) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M.nm || *M.failed) return;
  double q[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) q[c] = M.q[c * M.nm + m];
  write_sites(M, m, q);
}

// Everything between two force evaluations in one launch, MD_LANES lanes per
// molecule (a molecule's sites are consecutive, so its loads and stores of
// site positions and forces are coalesced): with `second`, the second half of
// step `step` (half kick, COM wrapped into the box, kinetic energies) and,
// from the last block to finish (a counter in failed[1]), that step's record
// as fixed-order sums over the molecules; with `first`, the first half of the
// next step (half kick, drift, NO_SQUISH rotation) and the molecule's sites
// for the next evaluation.  Both half kicks use the last evaluation's forces
// at the same orientation.  Every lane of a molecule advances the same body
// (the same arithmetic on the same values); its first lane stores it.
constexpr int MD_LANES = 32;
constexpr int MD_MOLS = 256 / MD_LANES;  // molecules per block
__global__ __launch_bounds__(256) void k_md_edge(MdDev M, int second, int step, int first) {
  __shared__ bool last;
  __shared__ double part[2][256];
  const int l = threadIdx.x % MD_LANES;
  const int m = blockIdx.x * MD_MOLS + threadIdx.x / MD_LANES;
  // once an evaluation outgrew its capacities the state stays put for the
  // rest of the run (the host repeats the run with larger lists)
  bool live = m < M.nm && !*M.failed;
  if (live) {
    if (const int why = evaluation_failure(M)) {
      if (m == 0 && l == 0) *M.failed = why;
      live = false;
    }
  }
  const long plane = (long)M.nm * M.ld, base = (long)m * M.ld;
  Body b;
  double F[3], tb[3];
  if (live) {
    b = load_body(M, m);
    // the molecular force and the lab-frame torque: each lane its sites, then
    // a fixed butterfly over the molecule's lanes
    double tl[3] = {0.0, 0.0, 0.0};
    F[0] = F[1] = F[2] = 0.0;
    for (int st = l; st < M.ns; st += MD_LANES) {
      double f[3], r[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        f[c] = M.fsite[c * plane + base + st];
        r[c] = M.site[c * plane + base + st];
        F[c] += f[c];
      }
      tl[0] += r[1] * f[2] - r[2] * f[1];
      tl[1] += r[2] * f[0] - r[0] * f[2];
      tl[2] += r[0] * f[1] - r[1] * f[0];
    }
#pragma unroll
    for (int d = MD_LANES / 2; d >= 1; d >>= 1)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        F[c] += __shfl_xor(F[c], d, MD_LANES);
        tl[c] += __shfl_xor(tl[c], d, MD_LANES);
      }
#pragma unroll
    for (int c = 0; c < 3; ++c) F[c] += M.fc2[c * M.nm + m] + M.fc3[c * M.nm + m];
    double R[9];
    rotation_matrix(b.q, R);
#pragma unroll
    for (int a = 0; a < 3; ++a) tb[a] = R[a] * tl[0] + R[3 + a] * tl[1] + R[6 + a] * tl[2];  // R^T tau
  }
  if (second) {
    if (live) {
      half_kick(M, b, F, tb);
      if (M.box > 0.0)
#pragma unroll
        for (int c = 0; c < 3; ++c) b.x[c] -= M.box * round(b.x[c] / M.box);  // NINT: half away from zero
      if (l == 0) {
        M.ke[m] = 0.5 * M.mass * (b.v[0] * b.v[0] + b.v[1] * b.v[1] + b.v[2] * b.v[2]);
        M.ke[M.nm + m] =
            0.5 * (b.j[0] * b.j[0] / M.pmi[0] + b.j[1] * b.j[1] / M.pmi[1] + b.j[2] * b.j[2] / M.pmi[2]);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      __threadfence();  // this block's energies before its count
      last = atomicAdd(M.failed + 1, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (last) {
      // the step's record: fixed-order sums over the molecules
      __threadfence();  // every other block's energies after their counts
      double kt = 0.0, kr = 0.0;
      for (int q = threadIdx.x; q < M.nm; q += 256) {
        kt += M.ke[q];
        kr += M.ke[M.nm + q];
      }
      part[0][threadIdx.x] = kt;
      part[1][threadIdx.x] = kr;
      __syncthreads();
      for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
          part[0][threadIdx.x] += part[0][threadIdx.x + w];
          part[1][threadIdx.x] += part[1][threadIdx.x + w];
        }
        __syncthreads();
      }
      if (threadIdx.x == 0) {
        const double e3 = M.en[0] + M.en[1];
        const double t = part[0][0] / 418.4, r = part[1][0] / 418.4;
        double* o = M.rec + 4 * (long)step;
        o[0] = e3;
        o[1] = t;
        o[2] = r;
        o[3] = e3 + t + r;
        M.failed[1] = 0;  // for the next launch
      }
    }
  }
  if (!live) return;
  if (first) {
    half_kick(M, b, F, tb);
#pragma unroll
    for (int c = 0; c < 3; ++c) b.x[c] += M.dt * b.v[c];
    rotate_rigid_body(M.dt, M.pmi, b.q, b.j);
    // the next evaluation's sites and zeroed site forces, a site per lane
    double R[9];
    rotation_matrix(b.q, R);
    for (int st = l; st < M.ld; st += MD_LANES) {
      double lab[3] = {0.0, 0.0, 0.0};
      if (st < M.ns) {
        const double* c = M.pc + 3 * st;
#pragma unroll
        for (int a = 0; a < 3; ++a) lab[a] = R[3 * a] * c[0] + R[3 * a + 1] * c[1] + R[3 * a + 2] * c[2];
      }
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        M.site[a * plane + base + st] = lab[a];
        M.fsite[a * plane + base + st] = 0.0;
      }
    }
  }
  if (l == 0) store_body(M, m, b);
}

}  // namespace ccpol3b

using namespace ccpol3b;

struct ccpol3b_md {
  ccpol3b_ctx* ctx = nullptr;  // borrowed
  MdDev M{};
  int cap_steps = 0;
  bool forces_current = false;
  // the force evaluation, recorded from the context into a HIP graph over the
  // driver's arrays (recorded again when the context's precision, coefficients
  // or tuning change, or a replay outgrew its list capacities)
  ccpol3b_graph* graph = nullptr;
  long graph_version = -1;  // context_version() when the graph was recorded
  double headroom = 0.25;
  double* saved = nullptr;  // [13][nm]: com, q, vel, jb at the start of a run
};

namespace {

template <typename P>
int md_alloc(P*& p, size_t n) {
  if (hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(P)) != hipSuccess) {
    p = nullptr;
    return CCPOL3B_ENOMEM;
  }
  return CCPOL3B_OK;
}

template <typename P>
void md_free(P*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// Records the force evaluation over the driver's arrays (sites must be current);
// the recording's own checked evaluation leaves valid forces behind.
int record_forces(ccpol3b_md* md, long version) {
  MdDev& M = md->M;
  if (md->graph) {
    ccpol3b_graph_destroy(md->graph);
    md->graph = nullptr;
  }
  const long plane = (long)M.nm * M.ld;
  const int ms[2] = {0, M.nm};
  const double bl[1] = {M.box};
  int rc = ccpol3b_graph_create(md->ctx, 1, ms, bl, M.ld, M.com, M.com + M.nm, M.com + 2 * M.nm, M.site,
                                M.site + plane, M.site + 2 * plane, M.fsite, M.fsite + plane, M.fsite + 2 * plane,
                                M.fc2, M.fc3, M.en, nullptr, md->headroom, nullptr, &md->graph);
  if (rc) return rc;
  md->graph_version = version;
  M.ftot = graph_device_totals(md->graph, &M.fcap3);
  return M.ftot ? CCPOL3B_OK : CCPOL3B_EHIP;
}

int launch_sites(const MdDev& M) {
  hipLaunchKernelGGL(k_md_sites, dim3((M.nm + 255) / 256), dim3(256), 0, nullptr, M);
  return hipGetLastError() == hipSuccess ? CCPOL3B_OK : CCPOL3B_EHIP;
}

// The arrays of the state, in the order they are saved.
void state_arrays(MdDev& M, double* a[4], size_t n[4]) {
  a[0] = M.com; n[0] = 3 * (size_t)M.nm;
  a[1] = M.q; n[1] = 4 * (size_t)M.nm;
  a[2] = M.vel; n[2] = 3 * (size_t)M.nm;
  a[3] = M.jb; n[3] = 3 * (size_t)M.nm;
}

int copy_state(ccpol3b_md* md, bool save) {
  double* a[4];
  size_t n[4], off = 0;
  state_arrays(md->M, a, n);
  for (int k = 0; k < 4; ++k) {
    double* dst = save ? md->saved + off : a[k];
    const double* src = save ? a[k] : md->saved + off;
    if (hipMemcpyAsync(dst, src, n[k] * sizeof(double), hipMemcpyDeviceToDevice, nullptr) != hipSuccess)
      return CCPOL3B_EHIP;
    off += n[k];
  }
  return CCPOL3B_OK;
}

}  // namespace

extern "C" {

void ccpol3b_md_destroy(ccpol3b_md* md) {
  if (!md) return;
  if (md->graph) ccpol3b_graph_destroy(md->graph);
  MdDev& M = md->M;
  md_free(M.com); md_free(M.q); md_free(M.vel); md_free(M.jb); md_free(M.pc); md_free(M.site);
  md_free(M.fsite); md_free(M.fc2); md_free(M.fc3); md_free(M.en); md_free(M.ke); md_free(M.rec);
  md_free(M.failed); md_free(md->saved);
  delete md;
}

int ccpol3b_md_create(ccpol3b_ctx* ctx, int nm, double box, int ns, const double* pcoord, double mass,
                      const double* pmi, ccpol3b_md** out) {
  if (!ctx || !out || nm <= 0 || ns < 17 || !pcoord || !(mass > 0.0) || !pmi || !(pmi[0] > 0.0) ||
      !(pmi[1] > 0.0) || !(pmi[2] > 0.0))
    return CCPOL3B_EINVAL;
  *out = nullptr;
  ccpol3b_md* md = new ccpol3b_md();
  md->ctx = ctx;
  MdDev& M = md->M;
  M.nm = nm;
  M.ns = ns;
  M.ld = ns;
  M.box = box;
  M.mass = mass;
  for (int a = 0; a < 3; ++a) M.pmi[a] = pmi[a];
  const size_t n3 = (size_t)3 * nm, nsite = (size_t)3 * nm * M.ld;
  if (md_alloc(M.com, n3) || md_alloc(M.q, (size_t)4 * nm) || md_alloc(M.vel, n3) || md_alloc(M.jb, n3) ||
      md_alloc(M.pc, (size_t)3 * ns) || md_alloc(M.site, nsite) || md_alloc(M.fsite, nsite) ||
      md_alloc(M.fc2, n3) || md_alloc(M.fc3, n3) || md_alloc(M.en, 3) || md_alloc(M.ke, (size_t)2 * nm) ||
      md_alloc(M.failed, 2) || md_alloc(md->saved, (size_t)13 * nm) ||
      hipMemcpy(M.pc, pcoord, sizeof(double) * 3 * ns, hipMemcpyHostToDevice) != hipSuccess) {
    ccpol3b_md_destroy(md);
    return CCPOL3B_ENOMEM;
  }
  // an inactive term leaves its COM force untouched (as the reference does),
  // and the kicks always add both: they start at zero
  if (hipMemset(M.fc2, 0, sizeof(double) * n3) != hipSuccess ||
      hipMemset(M.fc3, 0, sizeof(double) * n3) != hipSuccess) {
    ccpol3b_md_destroy(md);
    return CCPOL3B_EHIP;
  }
  *out = md;
  return CCPOL3B_OK;
}

// host arrays: com, vel, jbody [nm][3]; quat [nm][4]
int ccpol3b_md_set_state(ccpol3b_md* md, const double* com, const double* quat, const double* vel,
                         const double* jbody) {
  if (!md || !com || !quat || !vel || !jbody) return CCPOL3B_EINVAL;
  MdDev& M = md->M;
  const int nm = M.nm;
  std::vector<double> c(3 * nm), q(4 * nm), v(3 * nm), j(3 * nm);
  for (int m = 0; m < nm; ++m) {
    for (int a = 0; a < 3; ++a) {
      c[a * nm + m] = com[3 * m + a];
      v[a * nm + m] = vel[3 * m + a];
      j[a * nm + m] = jbody[3 * m + a];
    }
    for (int a = 0; a < 4; ++a) q[a * nm + m] = quat[4 * m + a];
  }
  if (hipMemcpy(M.com, c.data(), sizeof(double) * 3 * nm, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(M.q, q.data(), sizeof(double) * 4 * nm, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(M.vel, v.data(), sizeof(double) * 3 * nm, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(M.jb, j.data(), sizeof(double) * 3 * nm, hipMemcpyHostToDevice) != hipSuccess)
    return CCPOL3B_EHIP;
  md->forces_current = false;
  return CCPOL3B_OK;
}

int ccpol3b_md_get_state(ccpol3b_md* md, double* com, double* quat, double* vel, double* jbody) {
  if (!md) return CCPOL3B_EINVAL;
  MdDev& M = md->M;
  const int nm = M.nm;
  std::vector<double> c(3 * nm), q(4 * nm), v(3 * nm), j(3 * nm);
  if (hipMemcpy(c.data(), M.com, sizeof(double) * 3 * nm, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(q.data(), M.q, sizeof(double) * 4 * nm, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(v.data(), M.vel, sizeof(double) * 3 * nm, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(j.data(), M.jb, sizeof(double) * 3 * nm, hipMemcpyDeviceToHost) != hipSuccess)
    return CCPOL3B_EHIP;
  for (int m = 0; m < nm; ++m) {
    for (int a = 0; a < 3; ++a) {
      if (com) com[3 * m + a] = c[a * nm + m];
      if (vel) vel[3 * m + a] = v[a * nm + m];
      if (jbody) jbody[3 * m + a] = j[a * nm + m];
    }
    if (quat)
      for (int a = 0; a < 4; ++a) quat[4 * m + a] = q[a * nm + m];
  }
  return CCPOL3B_OK;
}

// nsteps steps of dt with no host synchronisation inside the loop: the force
// evaluation is a HIP graph replay.  The run ends with one wait; if a replay
// outgrew the recorded list capacities, the state was frozen from that step
// on, so the run is repeated from its start with the lists recorded again
// (more headroom).
int ccpol3b_md_run(ccpol3b_md* md, int nsteps, double dt, double* records) {
  if (!md || nsteps < 0 || !(dt > 0.0)) return CCPOL3B_EINVAL;
  MdDev& M = md->M;
  M.dt = dt;
  int rc;
  if (nsteps > md->cap_steps) {
    md_free(M.rec);
    if ((rc = md_alloc(M.rec, (size_t)4 * nsteps))) return rc;
    md->cap_steps = nsteps;
  }
  const long version = context_version(md->ctx);
  if ((rc = copy_state(md, true))) return rc;
  const dim3 grid((M.nm + MD_MOLS - 1) / MD_MOLS), block(256);
  for (int attempt = 0; attempt < 3; ++attempt) {
    if (hipMemsetAsync(M.failed, 0, 2 * sizeof(int), nullptr) != hipSuccess) return CCPOL3B_EHIP;
    if (!md->graph || md->graph_version != version) {
      // recording evaluates the forces at the current state
      if ((rc = launch_sites(M)) || (rc = record_forces(md, version))) return rc;
    } else if (!md->forces_current) {
      if ((rc = launch_sites(M)) || (rc = ccpol3b_graph_launch(md->graph, nullptr))) return rc;
    }
    // per step: the force evaluation, then one launch with the second half of
    // the step, its record and the first half of the next
    if (nsteps > 0) hipLaunchKernelGGL(k_md_edge, grid, block, 0, nullptr, M, 0, 0, 1);
    for (int s = 0; s < nsteps; ++s) {
      if ((rc = ccpol3b_graph_launch(md->graph, nullptr))) return rc;
      hipLaunchKernelGGL(k_md_edge, grid, block, 0, nullptr, M, 1, s, s + 1 < nsteps ? 1 : 0);
      if (hipGetLastError() != hipSuccess) return CCPOL3B_EHIP;
    }
    int failed = 0;
    if (hipMemcpy(&failed, M.failed, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return CCPOL3B_EHIP;
    // the last replay's outcome (no later stage kernel read it)
    const int last = ccpol3b_graph_status(md->graph);
    if (failed == FAIL_FORCE || last == CCPOL3B_EINVAL) return CCPOL3B_EINVAL;  // larger lists do not help
    if (!failed && last != CCPOL3B_OK) failed = FAIL_LISTS;
    if (!failed) {
      md->forces_current = true;
      if (records && nsteps > 0 &&
          hipMemcpy(records, M.rec, sizeof(double) * 4 * nsteps, hipMemcpyDeviceToHost) != hipSuccess)
        return CCPOL3B_EHIP;
      return CCPOL3B_OK;
    }
    // repeat from the start with larger lists
    if ((rc = copy_state(md, false))) return rc;
    md->headroom *= 2.0;
    ccpol3b_graph_destroy(md->graph);
    md->graph = nullptr;
    md->forces_current = false;
  }
  return CCPOL3B_ECAPACITY;
}

}  // extern "C"
