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
// thread iterates a group's constraints in the host's order until an
// iteration moves nothing, with the reference's updates and tests:
// positions as rattle_c (rattle.F90:180-226), velocities as rattlep_v
// (rattle_mtk.F90:455-490), evaluated without contraction like the
// reference's build; divisions by quantities that stay constant through the
// iteration are multiplications by reciprocals taken once.
#include "pgk_device.h"


namespace {

// half kick v += dt/2 F/m and drift x += dt v of one coordinate; every
// kernel that moves atoms uses these, so the arithmetic is the same
__device__ __forceinline__ double kick1(double v, double f, double minv, double dt) {
  const double h = 0.5 * dt * PGK_MD_CONV * minv;
  return v + h * f;
}
__device__ __forceinline__ double drift1(double x, double v, double dt) { return x + dt * v; }

// With the Nose-Hoover chain on, the half kick is the thermostatted update
// v exp(-dt vxi1/2) + F/m bbv (moldyn_nhc.F90:1301-1302, 1589-1590); nhc
// points at the two factors {exp(-dt vxi1/2), bbv} that k_nhc_half keeps on
// the device, or is null (plain NVE).
// With streaming velocities (strm), the chain scales the velocity relative to
// the atom's streaming velocity vs and adds vs back: ((v - vs) exp(-dt
// vxi1/2) + F/m bbv) + vs (moldyn_nhc.F90:1280-1310, 1585-1598).
struct Kick {
  bool nhc, strm;
  double s2, bbv;
  double vs[3];
};
__device__ __forceinline__ Kick kick_of(const double *nhc) {
  Kick k{nhc != nullptr, false, 1.0, 0.0, {0.0, 0.0, 0.0}};
  if (nhc) {
    k.s2 = nhc[0];
    k.bbv = nhc[1];
  }
  return k;
}
__device__ __forceinline__ double kick_nhc(double v, double f, double minv, const Kick &k,
                                           int d) {
#pragma clang fp contract(off)
  if (k.strm) return ((v - k.vs[d]) * k.s2 + f * PGK_MD_CONV * minv * k.bbv) + k.vs[d];
  return v * k.s2 + f * PGK_MD_CONV * minv * k.bbv;
}
__device__ __forceinline__ double kickv(double v, double f, double minv, double dt,
                                        const Kick &k, int d) {
  return k.nhc ? kick_nhc(v, f, minv, k, d) : kick1(v, f, minv, dt);
}
// With the region Langevin thermostat on as well, the chain acts on the atoms
// outside every Langevin region only: an atom registered in a region
// (ltreg[i] >= 0) takes the plain kick, the region thermostat its own update.
// vstrm: the atoms' streaming velocities [3*natom] under the chain, or null.
__device__ __forceinline__ Kick kick_at(const Kick &k, const int32_t *ltreg,
                                        const double *vstrm, int i) {
  if (k.nhc && ltreg && ltreg[i] >= 0) return Kick{false, false, 1.0, 0.0, {0.0, 0.0, 0.0}};
  Kick r = k;
  if (k.nhc && vstrm) {
    r.strm = true;
    for (int d = 0; d < 3; d++) r.vs[d] = vstrm[3 * i + d];
  }
  return r;
}

// RESPA (moldyn.F90:1465-1771): the long force's half kick (dt_l) comes
// before the short one on the first inner step and after it on the last;
// force_l is null on the other inner steps and without RESPA
__global__ __launch_bounds__(256) void k_md_kick_drift(int n, double *__restrict__ pos,
                                                       double *__restrict__ pos_old,
                                                       double *__restrict__ vel,
                                                       const double *__restrict__ force,
                                                       const double *__restrict__ minv, double dt,
                                                       const double *__restrict__ nhc,
                                                       const double *__restrict__ force_l,
                                                       double dt_l,
                                                       const int32_t *__restrict__ ltreg,
                                                       const double *__restrict__ vstrm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Kick k = kick_at(kick_of(nhc), ltreg, vstrm, i);
  for (int d = 0; d < 3; d++) {
    double v0 = vel[3 * i + d];
    if (force_l) v0 = kick1(v0, force_l[3 * i + d], minv[i], dt_l);
    const double v = kickv(v0, force[3 * i + d], minv[i], dt, k, d);
    vel[3 * i + d] = v;
    pos_old[3 * i + d] = pos[3 * i + d];
    pos[3 * i + d] = drift1(pos[3 * i + d], v, dt);
  }
}

__global__ __launch_bounds__(256) void k_md_kick(int n, double *__restrict__ vel,
                                                 const double *__restrict__ force,
                                                 const double *__restrict__ minv, double dt,
                                                 const double *__restrict__ nhc,
                                                 const double *__restrict__ force_l, double dt_l,
                                                 const int32_t *__restrict__ ltreg,
                                                 const double *__restrict__ vstrm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Kick k = kick_at(kick_of(nhc), ltreg, vstrm, i);
  for (int d = 0; d < 3; d++) {
    double v = kickv(vel[3 * i + d], force[3 * i + d], minv[i], dt, k, d);
    if (force_l) v = kick1(v, force_l[3 * i + d], minv[i], dt_l);
    vel[3 * i + d] = v;
  }
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
  double dt, tol;            // tol: eps_rattle, a relative tolerance on |r|^2
  int maxit;
  int32_t *fail;           // set when a group does not converge
  const double *nhc;       // thermostat factors of the half kicks, or null
  const double *force_l;   // RESPA: long force of this inner step's extra kick, or null
  double dt_l;             // RESPA: its time step (the outer one)
  const int32_t *ltreg;    // chain + region Langevin: region of each atom (-1 none), or null
  const double *vstrm;     // chain + streaming slabs: streaming velocity of each atom, or null
};

// Triangle groups (three atoms held by three constraints: the rigid 3-site
// water of the reference's topologies) are iterated in registers; set-up
// labels their atoms so that the host's constraint order reads (0,1), (0,2),
// (1,2).  Any other group iterates on global memory.
__device__ __forceinline__ double3 ld3(const double *p, int i) {
  return make_double3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}
__device__ __forceinline__ void st3(double *p, int i, const double3 &v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}
__device__ __forceinline__ double dot_nc(const double3 &a, const double3 &b) {
#pragma clang fp contract(off)
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ double3 sub(const double3 &a, const double3 &b) {
  return make_double3(a.x - b.x, a.y - b.y, a.z - b.z);
}

// One rattle_c update of constraint (i, j) (rattle.F90:180-226):
//   delta = |Rij^2 - d^2| / d^2, nothing to do when delta < eps;
//   g = (Rij^2 - d^2) / (2 RR (1/mi + 1/mj) dt), RR = Rij . Rij_old;
//   x_i -= (1/mi g Rij_old) dt, x_j += (1/mj g Rij_old) dt,
//   v_i -= 1/mi g Rij_old,      v_j += 1/mj g Rij_old.
// rd2 = 1/d^2: the division by the constant d^2 is a multiplication by its
// reciprocal (within one rounding of the reference's quotient), so each
// update keeps a single division.  Returns true when the constraint is
// already satisfied.
__device__ __forceinline__ bool shake_one(double3 &pi, double3 &pj, double3 &vi, double3 &vj,
                                          const double3 &ro, double mi, double mj, double d2,
                                          double rd2, double eps, double dt) {
#pragma clang fp contract(off)
  const double3 r = sub(pi, pj);
  const double rr2 = dot_nc(r, r);
  const bool ok = fabs(rr2 - d2) * rd2 < eps;
  // branch-free: a satisfied constraint applies g = 0, which leaves every
  // coordinate as it is
  const double gq = (rr2 - d2) / (2.0 * dot_nc(r, ro) * (mi + mj) * dt);
  const double g = ok ? 0.0 : gq;
  const double3 gr = make_double3(g * ro.x, g * ro.y, g * ro.z);
  pi = make_double3(pi.x - mi * gr.x * dt, pi.y - mi * gr.y * dt, pi.z - mi * gr.z * dt);
  pj = make_double3(pj.x + mj * gr.x * dt, pj.y + mj * gr.y * dt, pj.z + mj * gr.z * dt);
  vi = make_double3(vi.x - mi * gr.x, vi.y - mi * gr.y, vi.z - mi * gr.z);
  vj = make_double3(vj.x + mj * gr.x, vj.y + mj * gr.y, vj.z + mj * gr.z);
  return ok;
}

// One rattlep_v update of constraint (i, j) (rattle_mtk.F90:455-490):
//   phi = (Rij.Vij)^2 / Rij^2 * Vij^2, nothing to do when phi < eps^2;
//   k = Rij.Vij / (d^2 (1/mi + 1/mj)), v_i -= 1/mi k Rij, v_j += 1/mj k Rij.
// Positions do not move in this stage, so Rij^2 and d^2 (1/mi + 1/mj) are
// constants and both divisions are multiplications by their reciprocals
// (rr2i, rdm), each within one rounding of the reference's quotient.
__device__ __forceinline__ bool rattle_one(const double3 &r, double rr2i, double3 &vi,
                                           double3 &vj, double mi, double mj, double rdm,
                                           double eps2) {
#pragma clang fp contract(off)
  const double3 v = sub(vi, vj);
  const double rv = dot_nc(r, v);
  const bool ok = rv * rv * rr2i * dot_nc(v, v) < eps2;
  const double k = ok ? 0.0 : rv * rdm;   // branch-free, as in shake_one
  const double3 kr = make_double3(k * r.x, k * r.y, k * r.z);
  vi = make_double3(vi.x - mi * kr.x, vi.y - mi * kr.y, vi.z - mi * kr.z);
  vj = make_double3(vj.x + mj * kr.x, vj.y + mj * kr.y, vj.z + mj * kr.z);
  return ok;
}

// a triangle group's position stage in registers: constraints (0,1), (0,2),
// (1,2) in that order until an iteration moves nothing
__device__ __forceinline__ bool shake_triangle(const ConsArgs &a, int c0, int i0, int i1, int i2,
                                               double3 &p0, double3 &p1, double3 &p2,
                                               double3 &v0, double3 &v1, double3 &v2,
                                               const double3 &o0, const double3 &o1,
                                               const double3 &o2) {
  const double3 r01 = sub(o0, o1), r02 = sub(o0, o2), r12 = sub(o1, o2);
  const double m0 = a.minv[i0], m1 = a.minv[i1], m2 = a.minv[i2];
  const double d01 = a.d2[c0], d02 = a.d2[c0 + 1], d12 = a.d2[c0 + 2];
  const double e01 = 1.0 / d01, e02 = 1.0 / d02, e12 = 1.0 / d12;
  bool done = false;
  for (int it = 0; it < a.maxit && !done; it++) {
    done = shake_one(p0, p1, v0, v1, r01, m0, m1, d01, e01, a.tol, a.dt);
    done = shake_one(p0, p2, v0, v2, r02, m0, m2, d02, e02, a.tol, a.dt) && done;
    done = shake_one(p1, p2, v1, v2, r12, m1, m2, d12, e12, a.tol, a.dt) && done;
  }
  return done;
}

// a triangle group's velocity stage in registers
__device__ __forceinline__ bool rattle_triangle(const ConsArgs &a, int c0, int i0, int i1,
                                                int i2, const double3 &p0, const double3 &p1,
                                                const double3 &p2, double3 &v0, double3 &v1,
                                                double3 &v2) {
#pragma clang fp contract(off)
  const double m0 = a.minv[i0], m1 = a.minv[i1], m2 = a.minv[i2];
  const double eps2 = a.tol * a.tol;
  // positions are fixed here, so Rij and Rij^2 are the same every iteration
  const double3 r01 = sub(p0, p1), r02 = sub(p0, p2), r12 = sub(p1, p2);
  const double q01 = 1.0 / dot_nc(r01, r01), q02 = 1.0 / dot_nc(r02, r02),
               q12 = 1.0 / dot_nc(r12, r12);
  const double dm01 = 1.0 / (a.d2[c0] * (m0 + m1)), dm02 = 1.0 / (a.d2[c0 + 1] * (m0 + m2)),
               dm12 = 1.0 / (a.d2[c0 + 2] * (m1 + m2));
  bool done = false;
  for (int it = 0; it < a.maxit && !done; it++) {
    done = rattle_one(r01, q01, v0, v1, m0, m1, dm01, eps2);
    done = rattle_one(r02, q02, v0, v2, m0, m2, dm02, eps2) && done;
    done = rattle_one(r12, q12, v1, v2, m1, m2, dm12, eps2) && done;
  }
  return done;
}

// kick and drift of the three atoms of a triangle group, in registers
__device__ __forceinline__ void kick_drift3(const double *force, const double *minv, int i,
                                            double dt, const Kick &k, double3 &p, double3 &v,
                                            double3 &o, const double *force_l = nullptr,
                                            double dt_l = 0.0) {
  const double m = minv[i];
  o = p;
  if (force_l) {
    v.x = kick1(v.x, force_l[3 * i], m, dt_l);
    v.y = kick1(v.y, force_l[3 * i + 1], m, dt_l);
    v.z = kick1(v.z, force_l[3 * i + 2], m, dt_l);
  }
  v.x = kickv(v.x, force[3 * i], m, dt, k, 0);
  v.y = kickv(v.y, force[3 * i + 1], m, dt, k, 1);
  v.z = kickv(v.z, force[3 * i + 2], m, dt, k, 2);
  p.x = drift1(p.x, v.x, dt);
  p.y = drift1(p.y, v.y, dt);
  p.z = drift1(p.z, v.z, dt);
}
__device__ __forceinline__ void kick3(const double *force, const double *minv, int i, double dt,
                                      const Kick &k, double3 &v,
                                      const double *force_l = nullptr, double dt_l = 0.0) {
  const double m = minv[i];
  v.x = kickv(v.x, force[3 * i], m, dt, k, 0);
  v.y = kickv(v.y, force[3 * i + 1], m, dt, k, 1);
  v.z = kickv(v.z, force[3 * i + 2], m, dt, k, 2);
  if (force_l) {
    v.x = kick1(v.x, force_l[3 * i], m, dt_l);
    v.y = kick1(v.y, force_l[3 * i + 1], m, dt_l);
    v.z = kick1(v.z, force_l[3 * i + 2], m, dt_l);
  }
}

// When every group is a triangle: the first half step (kick, drift, position
// constraints) in one pass, one thread per group, then one thread per atom
// outside every group.  Same arithmetic as k_md_kick_drift + k_shake.
__global__ __launch_bounds__(64) void k_md_first_half(ConsArgs a, const double *__restrict__ force,
                                                      const int32_t *__restrict__ free_atoms,
                                                      int nfree) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const Kick k = kick_of(a.nhc);
  if (g < a.ngroup) {
    const int c0 = a.gstart[g];
    const int i0 = a.cij[2 * c0], i1 = a.cij[2 * c0 + 1], i2 = a.cij[2 * c0 + 3];
    double3 p0 = ld3(a.pos, i0), p1 = ld3(a.pos, i1), p2 = ld3(a.pos, i2);
    double3 v0 = ld3(a.vel, i0), v1 = ld3(a.vel, i1), v2 = ld3(a.vel, i2);
    double3 o0, o1, o2;
    kick_drift3(force, a.minv, i0, a.dt, kick_at(k, a.ltreg, a.vstrm, i0), p0,
                v0, o0, a.force_l, a.dt_l);
    kick_drift3(force, a.minv, i1, a.dt, kick_at(k, a.ltreg, a.vstrm, i1), p1,
                v1, o1, a.force_l, a.dt_l);
    kick_drift3(force, a.minv, i2, a.dt, kick_at(k, a.ltreg, a.vstrm, i2), p2,
                v2, o2, a.force_l, a.dt_l);
    const bool done = shake_triangle(a, c0, i0, i1, i2, p0, p1, p2, v0, v1, v2, o0, o1, o2);
    st3(a.pos, i0, p0);
    st3(a.pos, i1, p1);
    st3(a.pos, i2, p2);
    st3(a.vel, i0, v0);
    st3(a.vel, i1, v1);
    st3(a.vel, i2, v2);
    if (!done) atomicOr(a.fail, 1);
  } else if (g < a.ngroup + nfree) {
    const int i = free_atoms[g - a.ngroup];
    double3 p = ld3(a.pos, i), v = ld3(a.vel, i), o;
    kick_drift3(force, a.minv, i, a.dt, kick_at(k, a.ltreg, a.vstrm, i), p,
                v, o, a.force_l, a.dt_l);
    st3(a.pos, i, p);
    st3(a.vel, i, v);
  }
}

// the second half step (kick, velocity constraints) in one pass
__global__ __launch_bounds__(64) void k_md_second_half(ConsArgs a,
                                                       const double *__restrict__ force,
                                                       const int32_t *__restrict__ free_atoms,
                                                       int nfree) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const Kick k = kick_of(a.nhc);
  if (g < a.ngroup) {
    const int c0 = a.gstart[g];
    const int i0 = a.cij[2 * c0], i1 = a.cij[2 * c0 + 1], i2 = a.cij[2 * c0 + 3];
    const double3 p0 = ld3(a.pos, i0), p1 = ld3(a.pos, i1), p2 = ld3(a.pos, i2);
    double3 v0 = ld3(a.vel, i0), v1 = ld3(a.vel, i1), v2 = ld3(a.vel, i2);
    kick3(force, a.minv, i0, a.dt, kick_at(k, a.ltreg, a.vstrm, i0), v0, a.force_l, a.dt_l);
    kick3(force, a.minv, i1, a.dt, kick_at(k, a.ltreg, a.vstrm, i1), v1, a.force_l, a.dt_l);
    kick3(force, a.minv, i2, a.dt, kick_at(k, a.ltreg, a.vstrm, i2), v2, a.force_l, a.dt_l);
    const bool done = rattle_triangle(a, c0, i0, i1, i2, p0, p1, p2, v0, v1, v2);
    st3(a.vel, i0, v0);
    st3(a.vel, i1, v1);
    st3(a.vel, i2, v2);
    if (!done) atomicOr(a.fail, 2);
  } else if (g < a.ngroup + nfree) {
    const int i = free_atoms[g - a.ngroup];
    double3 v = ld3(a.vel, i);
    kick3(force, a.minv, i, a.dt, kick_at(k, a.ltreg, a.vstrm, i), v, a.force_l, a.dt_l);
    st3(a.vel, i, v);
  }
}

// positions (with the matching velocity correction), one group per thread
__global__ __launch_bounds__(64) void k_shake(ConsArgs a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.ngroup) return;
  const int c0 = a.gstart[g], c1 = a.gstart[g + 1];
  bool done = false;
  if (a.small[g]) {
    const int i0 = a.cij[2 * c0], i1 = a.cij[2 * c0 + 1], i2 = a.cij[2 * c0 + 3];
    double3 p0 = ld3(a.pos, i0), p1 = ld3(a.pos, i1), p2 = ld3(a.pos, i2);
    double3 v0 = ld3(a.vel, i0), v1 = ld3(a.vel, i1), v2 = ld3(a.vel, i2);
    const double3 o0 = ld3(a.pos_old, i0), o1 = ld3(a.pos_old, i1), o2 = ld3(a.pos_old, i2);
    done = shake_triangle(a, c0, i0, i1, i2, p0, p1, p2, v0, v1, v2, o0, o1, o2);
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
        const double3 ro = sub(ld3(a.pos_old, i), ld3(a.pos_old, j));
        if (!shake_one(pi, pj, vi, vj, ro, a.minv[i], a.minv[j], a.d2[k], 1.0 / a.d2[k], a.tol,
                       a.dt)) {
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

// velocities, one group per thread
__global__ __launch_bounds__(64) void k_rattle_v(ConsArgs a) {
#pragma clang fp contract(off)
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.ngroup) return;
  const int c0 = a.gstart[g], c1 = a.gstart[g + 1];
  const double eps2 = a.tol * a.tol;
  bool done = false;
  if (a.small[g]) {
    const int i0 = a.cij[2 * c0], i1 = a.cij[2 * c0 + 1], i2 = a.cij[2 * c0 + 3];
    const double3 p0 = ld3(a.pos, i0), p1 = ld3(a.pos, i1), p2 = ld3(a.pos, i2);
    double3 v0 = ld3(a.vel, i0), v1 = ld3(a.vel, i1), v2 = ld3(a.vel, i2);
    done = rattle_triangle(a, c0, i0, i1, i2, p0, p1, p2, v0, v1, v2);
    st3(a.vel, i0, v0);
    st3(a.vel, i1, v1);
    st3(a.vel, i2, v2);
  } else {
    for (int it = 0; it < a.maxit && !done; it++) {
      done = true;
      for (int k = c0; k < c1; k++) {
        const int i = a.cij[2 * k], j = a.cij[2 * k + 1];
        double3 vi = ld3(a.vel, i), vj = ld3(a.vel, j);
        const double mi = a.minv[i], mj = a.minv[j];
        const double3 r = sub(ld3(a.pos, i), ld3(a.pos, j));
        if (!rattle_one(r, 1.0 / dot_nc(r, r), vi, vj, mi, mj, 1.0 / (a.d2[k] * (mi + mj)),
                        eps2)) {
          done = false;
          st3(a.vel, i, vi);
          st3(a.vel, j, vj);
        }
      }
    }
  }
  if (!done) atomicOr(a.fail, 2);
}

// kinetic energy records: per block 0.5 sum m v^2 (fixed-order tree), v
// relative to the streaming velocity when vstrm is given (calkin.F90:98,
// nhcint.F90:53-54); with skip, the atoms registered in a Langevin region
// are left out
__global__ __launch_bounds__(256) void k_md_ke(int n, const double *__restrict__ vel,
                                               const double *__restrict__ minv,
                                               double *__restrict__ rec,
                                               const int32_t *__restrict__ skip,
                                               const double *__restrict__ vstrm) {
  double ev[1] = {0.0};
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(skip && skip[i] >= 0)) {
    double v[3] = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
    if (vstrm)
      for (int d = 0; d < 3; d++) v[d] = v[d] - vstrm[3 * i + d];
    const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    ev[0] = 0.5 * v2 / minv[i] / PGK_MD_CONV;
  }
  block_record<1, 4>(ev, rec + (size_t)blockIdx.x * PGK_REC);
}


// ---- Nose-Hoover chain ----------------------------------------------------
#define PGK_NHC_MAX 10   // maxnhc (md_global.F90:58)

struct NhcArgs {
  int m, nyosh, next, nrec;
  double gkt, gnkt, dt;
  double w2[5], w4[5], w8[5];   // yscoeff dt / next / 2, / 4, / 8 (nhcint.F90:56-58)
};

// the velocity update of chain thermostats 1 <= i <= m-2 (0-based), whose
// neighbour i+1 damps it (nhcint.F90:86-92, 108-114)
__device__ __forceinline__ double nhc_mid(const double (&q)[PGK_NHC_MAX],
                                          const double (&v)[PGK_NHC_MAX], int i, double gkt,
                                          double w4, double w8) {
#pragma clang fp contract(off)
  const double g = (q[i - 1] * v[i - 1] * v[i - 1] - gkt) / q[i];
  const double e = exp(-1.0 * w8 * v[i + 1]);
  return v[i] * e * e + w4 * g * e;
}

// One nhcint call (nhcint.F90:8-140) in one block: kint = sum m v^2 from the
// kinetic-energy records (a fixed-order sum), then thread 0 integrates the
// chain over next x nyosh sub-steps and refreshes the factors of the next
// half kicks (moldyn_nhc.F90:1285-1298).  The arrays are sized for the
// longest chain and indexed with constants, so they stay in registers;
// links past the chain hold v = 0 and are never stored.  st: [3m+2] =
// xlogs, vlogs, qmass, exp(-dt vxi1/2), bbv.
__global__ __launch_bounds__(256) void k_nhc_half(NhcArgs p, const double *__restrict__ rec,
                                                  double *__restrict__ st) {
#pragma clang fp contract(off)
  double t = 0.0;
  for (int b = threadIdx.x; b < p.nrec; b += 256) t += rec[(size_t)b * PGK_REC];
  __shared__ double part[4];
  t = wave_sum(t);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double kint = 2.0 * (((part[0] + part[1]) + part[2]) + part[3]);   // records hold m v^2 / 2
  const int m = p.m;
  double x[PGK_NHC_MAX], v[PGK_NHC_MAX], q[PGK_NHC_MAX];
#pragma unroll
  for (int i = 0; i < PGK_NHC_MAX; i++) {
    x[i] = i < m ? st[i] : 0.0;
    v[i] = i < m ? st[m + i] : 0.0;
    q[i] = i < m ? st[2 * m + i] : 1.0;
  }
  const double g1 = (kint - p.gnkt) / q[0];
  for (int inc = 0; inc < p.next; inc++) {
    for (int iy = 0; iy < p.nyosh; iy++) {
      const double w2 = p.w2[iy], w4 = p.w4[iy], w8 = p.w8[iy];
      // last thermostat, then m-1 down to 2 (1-based)
#pragma unroll
      for (int i = 1; i < PGK_NHC_MAX; i++)
        if (i == m - 1) v[i] = v[i] + (q[i - 1] * v[i - 1] * v[i - 1] - p.gkt) / q[i] * w4;
#pragma unroll
      for (int i = PGK_NHC_MAX - 2; i >= 1; i--)
        if (i <= m - 2) v[i] = nhc_mid(q, v, i, p.gkt, w4, w8);
      // first thermostat, the positions, the first thermostat again with
      // the same damping factor
      const double e = exp(-1.0 * w8 * v[1]);
      v[0] = v[0] * e * e + w4 * g1 * e;
#pragma unroll
      for (int i = 0; i < PGK_NHC_MAX; i++) x[i] = x[i] + v[i] * w2;
      v[0] = v[0] * e * e + w4 * g1 * e;
      // 2 up to m-1, then the last thermostat
#pragma unroll
      for (int i = 1; i <= PGK_NHC_MAX - 2; i++)
        if (i <= m - 2) v[i] = nhc_mid(q, v, i, p.gkt, w4, w8);
#pragma unroll
      for (int i = 1; i < PGK_NHC_MAX; i++)
        if (i == m - 1) v[i] = v[i] + (q[i - 1] * v[i - 1] * v[i - 1] - p.gkt) / q[i] * w4;
    }
  }
#pragma unroll
  for (int i = 0; i < PGK_NHC_MAX; i++)
    if (i < m) {
      st[i] = x[i];
      st[m + i] = v[i];
    }
  // expcdt2 = exp(-dt vxi1/4)^2, bbv = exp(-dt vxi1/4) sinh(dt vxi1/4)/(dt vxi1/4) dt/2
  const double e2 = 1.0 / 6.0, e4 = e2 / 20.0, e6 = e4 / 42.0, e8 = e6 / 72.0;
  const double ex4 = exp(-0.25 * p.dt * v[0]);
  const double a2 = (0.25 * p.dt * v[0]) * (0.25 * p.dt * v[0]);
  const double poly = 1.0 + a2 * (e2 + a2 * (e4 + a2 * (e6 + a2 * e8)));
  st[3 * m] = ex4 * ex4;
  st[3 * m + 1] = ex4 * poly * 0.5 * p.dt;
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
  a.nhc = c->md_nhc_m ? c->d_md_nhc + 3 * c->md_nhc_m : nullptr;
  a.force_l = nullptr;
  a.dt_l = 0.0;
  a.ltreg = (c->md_nhc_m && c->md_lt_nreg) ? c->d_lt_reg : nullptr;
  a.vstrm = (c->md_nhc_m && c->md_lt_nstrm) ? c->d_md_vstrm : nullptr;
  return a;
}

}  // namespace

// ---- launchers used by the C ABI (pgk_api.hip) ----------------------------
// kick (force f over dt, after the long force fl over dtl when given), drift
// over dt and the position constraints at dt
int pgk_md_kick_drift(pgk_ctx *c, const double *f, double dt, const double *fl, double dtl) {
  const int n = c->natom;
  ConsArgs a = cons_args(c);
  a.dt = dt;
  a.force_l = fl;
  a.dt_l = dtl;
  if (c->md_all_tri) {
    const int nt = c->md_ngroup + c->md_nfree;
    k_md_first_half<<<pgk_grid_blocks(nt, 64), 64, 0, c->stream>>>(a, f, c->d_md_free,
                                                                   c->md_nfree);
    HIP_CHECK(hipGetLastError());
    return pgk_md_strm_assign(c);   // at the new positions (moldyn_nhc.F90:1412)
  }
  k_md_kick_drift<<<pgk_grid_blocks(n, 256), 256, 0, c->stream>>>(
      n, c->d_pos, c->d_md_pos_old, c->d_md_vel, f, c->d_md_minv, dt, a.nhc, fl, dtl, a.ltreg,
      a.vstrm);
  HIP_CHECK(hipGetLastError());
  if (c->md_ngroup > 0) {
    k_shake<<<pgk_grid_blocks(c->md_ngroup, 64), 64, 0, c->stream>>>(a);
    HIP_CHECK(hipGetLastError());
  }
  return pgk_md_strm_assign(c);
}

// kick (force f over dt, then the long force fl over dtl when given) and the
// velocity constraints
int pgk_md_kick(pgk_ctx *c, const double *f, double dt, const double *fl, double dtl) {
  const int n = c->natom;
  ConsArgs a = cons_args(c);
  a.dt = dt;
  a.force_l = fl;
  a.dt_l = dtl;
  if (c->md_all_tri) {
    const int nt = c->md_ngroup + c->md_nfree;
    k_md_second_half<<<pgk_grid_blocks(nt, 64), 64, 0, c->stream>>>(a, f, c->d_md_free,
                                                                    c->md_nfree);
    HIP_CHECK(hipGetLastError());
    return PGK_OK;
  }
  k_md_kick<<<pgk_grid_blocks(n, 256), 256, 0, c->stream>>>(n, c->d_md_vel, f, c->d_md_minv,
                                                            dt, a.nhc, fl, dtl, a.ltreg,
                                                            a.vstrm);
  HIP_CHECK(hipGetLastError());
  return pgk_md_rattle_v(c);
}

int pgk_md_rattle_v(pgk_ctx *c) {
  if (c->md_ngroup > 0) {
    k_rattle_v<<<pgk_grid_blocks(c->md_ngroup, 64), 64, 0, c->stream>>>(cons_args(c));
    HIP_CHECK(hipGetLastError());
  }
  return PGK_OK;
}

// queue the kinetic-energy records into rec (pgk_grid_blocks(natom, 256)
// records); the chain's records (nhc) leave out the atoms of Langevin regions
int pgk_md_ke(pgk_ctx *c, double *rec, bool nhc) {
  const int n = c->natom;
  const int32_t *skip = (nhc && c->md_lt_nreg) ? c->d_lt_reg : nullptr;
  const double *vs = c->md_lt_nstrm ? c->d_md_vstrm : nullptr;
  k_md_ke<<<pgk_grid_blocks(n, 256), 256, 0, c->stream>>>(n, c->d_md_vel, c->d_md_minv, rec,
                                                          skip, vs);
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

// one nhcint call on the kinetic-energy records already in d_md_kenhc
int pgk_md_nhc_half(pgk_ctx *c) {
  NhcArgs p;
  p.m = c->md_nhc_m;
  p.nyosh = c->md_nhc_nyosh;
  p.next = c->md_nhc_next;
  p.nrec = pgk_grid_blocks(c->natom, 256);
  p.gkt = c->md_nhc_gkt;
  p.gnkt = c->md_nhc_gnkt;
  p.dt = c->md_dt;
  for (int k = 0; k < 5; k++) {
    p.w2[k] = c->md_nhc_ys[k] * c->md_dt / (double)c->md_nhc_next / 2.0;
    p.w4[k] = p.w2[k] / 2.0;
    p.w8[k] = p.w4[k] / 2.0;
  }
  k_nhc_half<<<1, 256, 0, c->stream>>>(p, c->d_md_kenhc, c->d_md_nhc);
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}
