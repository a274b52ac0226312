// k_md.hip - a device-resident NVE step around the force path: velocity
// Verlet with RATTLE constraints (Andersen 1983: SHAKE on positions, then
// the velocity stage), as the reference's moldyn/rattle pair does for rigid
// water (moldyn.F90:1465-1771, rattle.F90:6-272), kept on the GPU so a step
// needs no host round trip.
//
// Units: positions in A, forces in eref/A (the library's), masses in amu,
// velocities in A/fs, time in fs.  a = F/m * PGK_MD_CONV A/fs^2 and
// KE = 0.5 m v^2 / PGK_MD_CONV eref, with PGK_MD_CONV = eref / (amu A^2/fs^2)
// = 6.9511042e-21 J / 1.66053907e-17 J.
//
// Constraints are grouped by molecule (CSR over the constraint list); one
// thread iterates a group's constraints until every |r|^2 is within
// tol * d^2 (positions) or |r.v| <= tol * d^2 / dt-scale (velocities).
#include "pgk_device.h"

#define PGK_MD_CONV (6.9511042e-21 / 1.66053907e-17)

namespace {

__global__ __launch_bounds__(256) void k_md_kick_drift(int n, double *__restrict__ pos,
                                                       double *__restrict__ pos_old,
                                                       double *__restrict__ vel,
                                                       const double *__restrict__ force,
                                                       const double *__restrict__ minv, double dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double h = 0.5 * dt * PGK_MD_CONV * minv[i];
  for (int d = 0; d < 3; d++) {
    const double v = vel[3 * i + d] + h * force[3 * i + d];
    vel[3 * i + d] = v;
    pos_old[3 * i + d] = pos[3 * i + d];
    pos[3 * i + d] = pos[3 * i + d] + dt * v;
  }
}

__global__ __launch_bounds__(256) void k_md_kick(int n, double *__restrict__ vel,
                                                 const double *__restrict__ force,
                                                 const double *__restrict__ minv, double dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double h = 0.5 * dt * PGK_MD_CONV * minv[i];
  for (int d = 0; d < 3; d++) vel[3 * i + d] += h * force[3 * i + d];
}

struct ConsArgs {
  int ngroup;
  const int32_t *gstart;   // [ngroup+1] into the constraint list
  const int32_t *cij;      // [2*ncons] 0-based atoms
  const int32_t *small;    // [ngroup] 1: triangle group (constraints (0,1) (0,2) (1,2))
  const double *d2;        // [ncons] squared lengths
  const double *minv;      // [natom]
  double *pos, *vel;
  const double *pos_old;
  double dt, tol;
  int maxit;
  int32_t *fail;           // set when a group does not converge
};

// Triangle groups (three atoms held by three constraints: the rigid 3-site
// water of the reference's topologies) are iterated in registers with the
// constraints reordered at set-up to (0,1), (0,2), (1,2) of the group's
// atoms; any other group iterates on global memory.
__device__ __forceinline__ double3 ld3(const double *p, int i) {
  return make_double3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}
__device__ __forceinline__ void st3(double *p, int i, const double3 &v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}
__device__ __forceinline__ double dot(const double3 &a, const double3 &b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ double3 sub(const double3 &a, const double3 &b) {
  return make_double3(a.x - b.x, a.y - b.y, a.z - b.z);
}
__device__ __forceinline__ void axpy(double3 &x, double s, const double3 &v) {
  x.x += s * v.x;
  x.y += s * v.y;
  x.z += s * v.z;
}

// one SHAKE update of constraint (i, j) with squared length d2
__device__ __forceinline__ bool shake_one(double3 &pi, double3 &pj, double3 &vi, double3 &vj,
                                          const double3 &oi, const double3 &oj, double mi,
                                          double mj, double d2, double tol, double dt) {
  const double3 r = sub(pi, pj);
  const double diff = d2 - dot(r, r);
  if (fabs(diff) <= 2.0 * tol * d2) return true;
  const double3 ro = sub(oi, oj);
  const double gg = diff / (2.0 * (mi + mj) * dot(r, ro));
  axpy(pi, gg * mi, ro);
  axpy(pj, -gg * mj, ro);
  axpy(vi, gg * mi / dt, ro);
  axpy(vj, -gg * mj / dt, ro);
  return false;
}

// one RATTLE velocity update of constraint (i, j)
__device__ __forceinline__ bool rattle_one(const double3 &pi, const double3 &pj, double3 &vi,
                                           double3 &vj, double mi, double mj, double d2,
                                           double tol) {
  const double3 r = sub(pi, pj);
  const double rv = dot(r, sub(vi, vj));
  if (fabs(rv) <= tol * d2) return true;
  const double kk = rv / (d2 * (mi + mj));
  axpy(vi, -kk * mi, r);
  axpy(vj, kk * mj, r);
  return false;
}

// SHAKE on the positions (with the matching velocity correction): displace
// along the old bond vector, g = (d^2 - r.r) / (2 (1/mi + 1/mj) r.r_old)
__global__ __launch_bounds__(256) void k_shake(ConsArgs a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.ngroup) return;
  const int c0 = a.gstart[g], c1 = a.gstart[g + 1];
  bool done = false;
  if (a.small[g]) {
    const int i0 = a.cij[2 * c0], i1 = a.cij[2 * c0 + 1], i2 = a.cij[2 * c0 + 3];
    double3 p0 = ld3(a.pos, i0), p1 = ld3(a.pos, i1), p2 = ld3(a.pos, i2);
    double3 v0 = ld3(a.vel, i0), v1 = ld3(a.vel, i1), v2 = ld3(a.vel, i2);
    const double3 o0 = ld3(a.pos_old, i0), o1 = ld3(a.pos_old, i1), o2 = ld3(a.pos_old, i2);
    const double m0 = a.minv[i0], m1 = a.minv[i1], m2 = a.minv[i2];
    const double d01 = a.d2[c0], d02 = a.d2[c0 + 1], d12 = a.d2[c0 + 2];
    for (int it = 0; it < a.maxit && !done; it++) {
      done = shake_one(p0, p1, v0, v1, o0, o1, m0, m1, d01, a.tol, a.dt);
      done = shake_one(p0, p2, v0, v2, o0, o2, m0, m2, d02, a.tol, a.dt) && done;
      done = shake_one(p1, p2, v1, v2, o1, o2, m1, m2, d12, a.tol, a.dt) && done;
    }
    st3(a.pos, i0, p0);
    st3(a.pos, i1, p1);
    st3(a.pos, i2, p2);
    st3(a.vel, i0, v0);
    st3(a.vel, i1, v1);
    st3(a.vel, i2, v2);
  } else {
    for (int it = 0; it < a.maxit && !done; it++) {
      done = true;
      for (int k = c0; k < c1; k++) {
        const int i = a.cij[2 * k], j = a.cij[2 * k + 1];
        double3 pi = ld3(a.pos, i), pj = ld3(a.pos, j), vi = ld3(a.vel, i), vj = ld3(a.vel, j);
        if (!shake_one(pi, pj, vi, vj, ld3(a.pos_old, i), ld3(a.pos_old, j), a.minv[i],
                       a.minv[j], a.d2[k], a.tol, a.dt)) {
          done = false;
          st3(a.pos, i, pi);
          st3(a.pos, j, pj);
          st3(a.vel, i, vi);
          st3(a.vel, j, vj);
        }
      }
    }
  }
  if (!done) atomicOr(a.fail, 1);
}

// RATTLE velocity stage: remove the relative velocity along each bond,
// k = r.(vi - vj) / (d^2 (1/mi + 1/mj))
__global__ __launch_bounds__(256) void k_rattle_v(ConsArgs a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.ngroup) return;
  const int c0 = a.gstart[g], c1 = a.gstart[g + 1];
  bool done = false;
  if (a.small[g]) {
    const int i0 = a.cij[2 * c0], i1 = a.cij[2 * c0 + 1], i2 = a.cij[2 * c0 + 3];
    const double3 p0 = ld3(a.pos, i0), p1 = ld3(a.pos, i1), p2 = ld3(a.pos, i2);
    double3 v0 = ld3(a.vel, i0), v1 = ld3(a.vel, i1), v2 = ld3(a.vel, i2);
    const double m0 = a.minv[i0], m1 = a.minv[i1], m2 = a.minv[i2];
    const double d01 = a.d2[c0], d02 = a.d2[c0 + 1], d12 = a.d2[c0 + 2];
    for (int it = 0; it < a.maxit && !done; it++) {
      done = rattle_one(p0, p1, v0, v1, m0, m1, d01, a.tol);
      done = rattle_one(p0, p2, v0, v2, m0, m2, d02, a.tol) && done;
      done = rattle_one(p1, p2, v1, v2, m1, m2, d12, a.tol) && done;
    }
    st3(a.vel, i0, v0);
    st3(a.vel, i1, v1);
    st3(a.vel, i2, v2);
  } else {
    for (int it = 0; it < a.maxit && !done; it++) {
      done = true;
      for (int k = c0; k < c1; k++) {
        const int i = a.cij[2 * k], j = a.cij[2 * k + 1];
        double3 vi = ld3(a.vel, i), vj = ld3(a.vel, j);
        if (!rattle_one(ld3(a.pos, i), ld3(a.pos, j), vi, vj, a.minv[i], a.minv[j], a.d2[k],
                        a.tol)) {
          done = false;
          st3(a.vel, i, vi);
          st3(a.vel, j, vj);
        }
      }
    }
  }
  if (!done) atomicOr(a.fail, 2);
}

// kinetic energy records: per block 0.5 sum m v^2 (fixed-order tree)
__global__ __launch_bounds__(256) void k_md_ke(int n, const double *__restrict__ vel,
                                               const double *__restrict__ minv,
                                               double *__restrict__ rec) {
  double ev[1] = {0.0};
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double v2 = vel[3 * i] * vel[3 * i] + vel[3 * i + 1] * vel[3 * i + 1] +
                      vel[3 * i + 2] * vel[3 * i + 2];
    ev[0] = 0.5 * v2 / minv[i] / PGK_MD_CONV;
  }
  block_record<1, 4>(ev, rec + (size_t)blockIdx.x * PGK_REC);
}

ConsArgs cons_args(pgk_ctx *c) {
  ConsArgs a;
  a.ngroup = c->md_ngroup;
  a.gstart = c->d_md_gstart;
  a.cij = c->d_md_cij;
  a.small = c->d_md_small;
  a.d2 = c->d_md_d2;
  a.minv = c->d_md_minv;
  a.pos = c->d_pos;
  a.vel = c->d_md_vel;
  a.pos_old = c->d_md_pos_old;
  a.dt = c->md_dt;
  a.tol = c->md_tol;
  a.maxit = 500;
  a.fail = c->d_md_fail;
  return a;
}

}  // namespace

// ---- launchers used by the C ABI (pgk_api.hip) ----------------------------
int pgk_md_kick_drift(pgk_ctx *c) {
  const int n = c->natom;
  k_md_kick_drift<<<pgk_grid_blocks(n, 256), 256, 0, c->stream>>>(
      n, c->d_pos, c->d_md_pos_old, c->d_md_vel, c->d_force, c->d_md_minv, c->md_dt);
  HIP_CHECK(hipGetLastError());
  if (c->md_ngroup > 0) {
    k_shake<<<pgk_grid_blocks(c->md_ngroup, 256), 256, 0, c->stream>>>(cons_args(c));
    HIP_CHECK(hipGetLastError());
  }
  return PGK_OK;
}

int pgk_md_kick(pgk_ctx *c) {
  const int n = c->natom;
  k_md_kick<<<pgk_grid_blocks(n, 256), 256, 0, c->stream>>>(n, c->d_md_vel, c->d_force,
                                                            c->d_md_minv, c->md_dt);
  HIP_CHECK(hipGetLastError());
  return pgk_md_rattle_v(c);
}

int pgk_md_rattle_v(pgk_ctx *c) {
  if (c->md_ngroup > 0) {
    k_rattle_v<<<pgk_grid_blocks(c->md_ngroup, 256), 256, 0, c->stream>>>(cons_args(c));
    HIP_CHECK(hipGetLastError());
  }
  return PGK_OK;
}

// queue the kinetic-energy records into rec (pgk_grid_blocks(natom, 256) records)
int pgk_md_ke(pgk_ctx *c, double *rec) {
  const int n = c->natom;
  k_md_ke<<<pgk_grid_blocks(n, 256), 256, 0, c->stream>>>(n, c->d_md_vel, c->d_md_minv, rec);
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}
