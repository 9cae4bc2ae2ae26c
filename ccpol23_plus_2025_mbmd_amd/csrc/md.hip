// Rigid-water NVE driver on the device: velocity Verlet for the centres of
// mass and the NO_SQUISH symplectic rotation for the orientations, as
// nve_1 (evolve.f90:621-1062) and nosquish (evolve.f90:8107-8204) integrate
// them, with the forces of the three-body term evaluated by this library each
// step.  State per molecule: COM (A), quaternion, COM velocity (A/ps) and
// body-frame angular momentum (amu A^2/ps); forces in 10 J mol^-1 A^-1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/ccpol3b.h"
#include "ccpol3b_internal.h"

namespace ccpol3b {

struct MdDev {
  int nm, ns, ld;
  double dt, totm, pmi[3], box;
  double *com;    // [3][nm]
  double *q;      // [4][nm]
  double *vel;    // [3][nm]
  double *jb;     // [3][nm]
  double *pc;     // [ns][3] principal-frame sites
  double *site;   // [3][nm][ld] lab-frame offsets
  double *fsite;  // [3][nm][ld]
  double *fc2, *fc3;  // [3][nm]
  double *en;     // [3] qfs2, qfs3, vir increment
  double *ke;     // [2][nm] translational, rotational kinetic energy per molecule
  double *rec;    // [nsteps][4] E3, Kt, Kr, E3 + Kt + Kr (kcal/mol)
};

// rott(1..9) of quat_to_rot (molecular_sites.f90:95-111), row major
__device__ __forceinline__ void md_rot(double q0, double q1, double q2, double q3, double* r) {
  r[0] = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3;
  r[1] = 2.0 * (q1 * q2 - q0 * q3);
  r[2] = 2.0 * (q1 * q3 + q0 * q2);
  r[3] = 2.0 * (q1 * q2 + q0 * q3);
  r[4] = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3;
  r[5] = 2.0 * (q2 * q3 - q0 * q1);
  r[6] = 2.0 * (q1 * q3 - q0 * q2);
  r[7] = 2.0 * (q2 * q3 + q0 * q1);
  r[8] = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3;
}

// Lab-frame site offsets R(q) pc (transform, iway = 1) and zeroed site forces.
__global__ void k_md_sites(MdDev M) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)M.nm * M.ld) return;
  const int m = (int)(t / M.ld), s = (int)(t - (long)m * M.ld);
  const long plane = (long)M.nm * M.ld;
  double x = 0.0, y = 0.0, z = 0.0;
  if (s < M.ns) {
    double r[9];
    md_rot(M.q[m], M.q[M.nm + m], M.q[2 * M.nm + m], M.q[3 * M.nm + m], r);
    const double* p = M.pc + 3 * s;
    x = p[0] * r[0] + p[1] * r[1] + p[2] * r[2];
    y = p[0] * r[3] + p[1] * r[4] + p[2] * r[5];
    z = p[0] * r[6] + p[1] * r[7] + p[2] * r[8];
  }
  M.site[t] = x;
  M.site[plane + t] = y;
  M.site[2 * plane + t] = z;
  M.fsite[t] = 0.0;
  M.fsite[plane + t] = 0.0;
  M.fsite[2 * plane + t] = 0.0;
}

// nosquish (evolve.f90:8107-8204): 10 sub-steps of the z/2, y/2, x, y/2, z/2
// free-rotor splitting of the quaternion and its conjugate momentum.
__device__ void md_nosquish(double tstep, const double* pmi, double* qq, double* pp) {
  const int mrot = 10;
  const double ov4 = 0.25, ov8 = 0.125;
  const double tr = tstep / (double)mrot;
  for (int m = 0; m < mrot; ++m) {
    double cs, sn, z;
    double qn1[4], pq2[4], qn2[4], pq3[4], qn3[4], pq4[4];
    z = ov8 / pmi[2] * tr * (-pp[0] * qq[3] + pp[1] * qq[2] - pp[2] * qq[1] + pp[3] * qq[0]);
    cs = cos(z);
    sn = sin(z);
    qn1[0] = cs * qq[0] - sn * qq[3]; qn1[1] = cs * qq[1] + sn * qq[2];
    qn1[2] = cs * qq[2] - sn * qq[1]; qn1[3] = cs * qq[3] + sn * qq[0];
    pq2[0] = cs * pp[0] - sn * pp[3]; pq2[1] = cs * pp[1] + sn * pp[2];
    pq2[2] = cs * pp[2] - sn * pp[1]; pq2[3] = cs * pp[3] + sn * pp[0];
    z = ov8 / pmi[1] * tr * (-pq2[0] * qn1[2] - pq2[1] * qn1[3] + pq2[2] * qn1[0] + pq2[3] * qn1[1]);
    cs = cos(z);
    sn = sin(z);
    qn2[0] = cs * qn1[0] - sn * qn1[2]; qn2[1] = cs * qn1[1] - sn * qn1[3];
    qn2[2] = cs * qn1[2] + sn * qn1[0]; qn2[3] = cs * qn1[3] + sn * qn1[1];
    pq3[0] = cs * pq2[0] - sn * pq2[2]; pq3[1] = cs * pq2[1] - sn * pq2[3];
    pq3[2] = cs * pq2[2] + sn * pq2[0]; pq3[3] = cs * pq2[3] + sn * pq2[1];
    z = ov4 / pmi[0] * tr * (-pq3[0] * qn2[1] + pq3[1] * qn2[0] + pq3[2] * qn2[3] - pq3[3] * qn2[2]);
    cs = cos(z);
    sn = sin(z);
    qn3[0] = cs * qn2[0] - sn * qn2[1]; qn3[1] = cs * qn2[1] + sn * qn2[0];
    qn3[2] = cs * qn2[2] + sn * qn2[3]; qn3[3] = cs * qn2[3] - sn * qn2[2];
    pq4[0] = cs * pq3[0] - sn * pq3[1]; pq4[1] = cs * pq3[1] + sn * pq3[0];
    pq4[2] = cs * pq3[2] + sn * pq3[3]; pq4[3] = cs * pq3[3] - sn * pq3[2];
    z = ov8 / pmi[1] * tr * (-pq4[0] * qn3[2] - pq4[1] * qn3[3] + pq4[2] * qn3[0] + pq4[3] * qn3[1]);
    cs = cos(z);
    sn = sin(z);
    qn2[0] = cs * qn3[0] - sn * qn3[2]; qn2[1] = cs * qn3[1] - sn * qn3[3];
    qn2[2] = cs * qn3[2] + sn * qn3[0]; qn2[3] = cs * qn3[3] + sn * qn3[1];
    pq3[0] = cs * pq4[0] - sn * pq4[2]; pq3[1] = cs * pq4[1] - sn * pq4[3];
    pq3[2] = cs * pq4[2] + sn * pq4[0]; pq3[3] = cs * pq4[3] + sn * pq4[1];
    z = ov8 / pmi[2] * tr * (-pq3[0] * qn2[3] + pq3[1] * qn2[2] - pq3[2] * qn2[1] + pq3[3] * qn2[0]);
    cs = cos(z);
    sn = sin(z);
    qq[0] = cs * qn2[0] - sn * qn2[3]; qq[1] = cs * qn2[1] + sn * qn2[2];
    qq[2] = cs * qn2[2] - sn * qn2[1]; qq[3] = cs * qn2[3] + sn * qn2[0];
    pp[0] = cs * pq3[0] - sn * pq3[3]; pp[1] = cs * pq3[1] + sn * pq3[2];
    pp[2] = cs * pq3[2] - sn * pq3[1]; pp[3] = cs * pq3[3] + sn * pq3[0];
  }
}

// One half of the velocity-Verlet step for molecule m (evolve.f90:659-781
// for stage 1, :889-1017 for stage 2): molecular force and torque from the
// site and COM forces, half kicks of the COM velocity and of the quaternion
// momentum; stage 1 then drifts the COM and rotates with NO_SQUISH, stage 2
// wraps the COM into the box and records the kinetic energies.
__global__ void k_md_kick(MdDev M, int stage) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M.nm) return;
  const int nm = M.nm;
  const long plane = (long)nm * M.ld, b = (long)m * M.ld;
  double F[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
  for (int s = 0; s < M.ns; ++s) {
    const double fx = M.fsite[b + s], fy = M.fsite[plane + b + s], fz = M.fsite[2 * plane + b + s];
    const double x = M.site[b + s], y = M.site[plane + b + s], z = M.site[2 * plane + b + s];
    F[0] += fx;
    F[1] += fy;
    F[2] += fz;
    T[0] += y * fz - z * fy;
    T[1] += z * fx - x * fz;
    T[2] += x * fy - y * fx;
  }
  for (int c = 0; c < 3; ++c) F[c] = F[c] + M.fc2[c * nm + m] + M.fc3[c * nm + m];
  double q[4] = {M.q[m], M.q[nm + m], M.q[2 * nm + m], M.q[3 * nm + m]};
  double r[9];
  md_rot(q[0], q[1], q[2], q[3], r);
  const double tqx = T[0] * r[0] + T[1] * r[3] + T[2] * r[6];
  const double tqy = T[0] * r[1] + T[1] * r[4] + T[2] * r[7];
  const double tqz = T[0] * r[2] + T[1] * r[5] + T[2] * r[8];
  const double qt0 = 2.0 * (-q[1] * tqx - q[2] * tqy - q[3] * tqz);
  const double qt1 = 2.0 * (q[0] * tqx - q[3] * tqy + q[2] * tqz);
  const double qt2 = 2.0 * (q[3] * tqx + q[0] * tqy - q[1] * tqz);
  const double qt3 = 2.0 * (-q[2] * tqx + q[1] * tqy + q[0] * tqz);
  const double opx = M.jb[m], opy = M.jb[nm + m], opz = M.jb[2 * nm + m];
  double p[4];
  p[0] = 2.0 * (-q[1] * opx - q[2] * opy - q[3] * opz);
  p[1] = 2.0 * (q[0] * opx - q[3] * opy + q[2] * opz);
  p[2] = 2.0 * (q[3] * opx + q[0] * opy - q[1] * opz);
  p[3] = 2.0 * (-q[2] * opx + q[1] * opy + q[0] * opz);
  p[0] = p[0] + qt0 * 0.5 * M.dt;
  p[1] = p[1] + qt1 * 0.5 * M.dt;
  p[2] = p[2] + qt2 * 0.5 * M.dt;
  p[3] = p[3] + qt3 * 0.5 * M.dt;
  double v[3];
  for (int c = 0; c < 3; ++c) v[c] = M.vel[c * nm + m] + 0.5 * F[c] * M.dt / M.totm;
  double x[3] = {M.com[m], M.com[nm + m], M.com[2 * nm + m]};
  if (stage == 1) {
    for (int c = 0; c < 3; ++c) x[c] = x[c] + M.dt * v[c];
    md_nosquish(M.dt, M.pmi, q, p);
  } else if (M.box > 0.0) {
    for (int c = 0; c < 3; ++c) x[c] = x[c] - M.box * round(x[c] / M.box);
  }
  const double jx = 0.5 * (-q[1] * p[0] + q[0] * p[1] + q[3] * p[2] - q[2] * p[3]);
  const double jy = 0.5 * (-q[2] * p[0] - q[3] * p[1] + q[0] * p[2] + q[1] * p[3]);
  const double jz = 0.5 * (-q[3] * p[0] + q[2] * p[1] - q[1] * p[2] + q[0] * p[3]);
  for (int c = 0; c < 3; ++c) {
    M.com[c * nm + m] = x[c];
    M.vel[c * nm + m] = v[c];
  }
  for (int c = 0; c < 4; ++c) M.q[c * nm + m] = q[c];
  M.jb[m] = jx;
  M.jb[nm + m] = jy;
  M.jb[2 * nm + m] = jz;
  if (stage == 2) {
    M.ke[m] = 0.5 * M.totm * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);  // tkinetic_eng (utility.f90:164-179)
    M.ke[nm + m] = 0.5 * (jx * jx / M.pmi[0] + jy * jy / M.pmi[1] + jz * jz / M.pmi[2]);  // rkinetic_eng
  }
}

// Step record: E3 = qfs2 + qfs3, Kt, Kr (kcal/mol) and their sum; one block,
// fixed-order sums.
__global__ __launch_bounds__(256) void k_md_record(MdDev M, int step) {
  __shared__ double red[2][256];
  double kt = 0.0, kr = 0.0;
  for (int m = threadIdx.x; m < M.nm; m += 256) {
    kt += M.ke[m];
    kr += M.ke[M.nm + m];
  }
  red[0][threadIdx.x] = kt;
  red[1][threadIdx.x] = kr;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double e3 = M.en[0] + M.en[1];
    const double t = red[0][0] / 418.4, r = red[1][0] / 418.4;
    double* o = M.rec + 4 * (long)step;
    o[0] = e3;
    o[1] = t;
    o[2] = r;
    o[3] = e3 + t + r;
  }
}

}  // namespace ccpol3b

using namespace ccpol3b;

struct ccpol3b_md {
  ccpol3b_ctx* ctx = nullptr;  // borrowed
  MdDev M{};
  int cap_steps = 0;
  bool forces_current = false;
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

// site offsets, then the three-body forces at the current state
int md_forces(ccpol3b_md* md) {
  MdDev& M = md->M;
  const long ns_tot = (long)M.nm * M.ld;
  hipLaunchKernelGGL(k_md_sites, dim3((ns_tot + 255) / 256), dim3(256), 0, nullptr, M);
  if (hipGetLastError() != hipSuccess) return CCPOL3B_EHIP;
  const long plane = (long)M.nm * M.ld;
  int rc = ccpol3b_eval_device(md->ctx, M.nm, M.ld, M.com, M.com + M.nm, M.com + 2 * M.nm, M.site, M.site + plane,
                               M.site + 2 * plane, M.box, M.fsite, M.fsite + plane, M.fsite + 2 * plane, M.fc2,
                               M.fc3, M.en, nullptr);
  if (rc == CCPOL3B_OK) md->forces_current = true;
  return rc;
}

}  // namespace

extern "C" {

void ccpol3b_md_destroy(ccpol3b_md* md) {
  if (!md) return;
  MdDev& M = md->M;
  md_free(M.com); md_free(M.q); md_free(M.vel); md_free(M.jb); md_free(M.pc); md_free(M.site);
  md_free(M.fsite); md_free(M.fc2); md_free(M.fc3); md_free(M.en); md_free(M.ke); md_free(M.rec);
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
  M.totm = mass;
  for (int a = 0; a < 3; ++a) M.pmi[a] = pmi[a];
  const size_t n3 = (size_t)3 * nm, nsite = (size_t)3 * nm * M.ld;
  if (md_alloc(M.com, n3) || md_alloc(M.q, (size_t)4 * nm) || md_alloc(M.vel, n3) || md_alloc(M.jb, n3) ||
      md_alloc(M.pc, (size_t)3 * ns) || md_alloc(M.site, nsite) || md_alloc(M.fsite, nsite) ||
      md_alloc(M.fc2, n3) || md_alloc(M.fc3, n3) || md_alloc(M.en, 3) || md_alloc(M.ke, (size_t)2 * nm) ||
      hipMemcpy(M.pc, pcoord, sizeof(double) * 3 * ns, hipMemcpyHostToDevice) != hipSuccess) {
    ccpol3b_md_destroy(md);
    return CCPOL3B_ENOMEM;
  }
  // an inactive term leaves its COM force untouched (as the reference does), and
  // the kick always adds both: they start at zero
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
  if (!md->forces_current && (rc = md_forces(md))) return rc;
  const dim3 g((M.nm + 255) / 256), b(256);
  for (int s = 0; s < nsteps; ++s) {
    hipLaunchKernelGGL(k_md_kick, g, b, 0, nullptr, M, 1);
    if ((rc = md_forces(md))) return rc;
    hipLaunchKernelGGL(k_md_kick, g, b, 0, nullptr, M, 2);
    hipLaunchKernelGGL(k_md_record, dim3(1), dim3(256), 0, nullptr, M, s);
    if (hipGetLastError() != hipSuccess) return CCPOL3B_EHIP;
  }
  if (records && nsteps > 0 &&
      hipMemcpy(records, M.rec, sizeof(double) * 4 * nsteps, hipMemcpyDeviceToHost) != hipSuccess)
    return CCPOL3B_EHIP;
  if (hipDeviceSynchronize() != hipSuccess) return CCPOL3B_EHIP;
  return CCPOL3B_OK;
}

}  // extern "C"
