// k_hf_w.hip - the per-atom heat-flux arrays of calforce_hf over the listed
// pairs: per-atom energies pot_*_atm and the calvirihf tensors viri*t_atm.
//
// Restates the per-atom part of ewaldrp_hf (ewaldr_hf.F90:8-670) in the
// default build, as the pair pass of calforce_hf: for every listed pair with
// r <= rrcut the bare Coulomb energy qq/r and force qq/r^3 r_ij
// (ewaldr_hf.F90:336-343), for r^2 <= rcut^2 and INTTYPE_VDW the LJ energy and
// force; half of each energy goes to each atom, class ELIN/LJIN for pairs
// inside one molecule and ELC/VDW otherwise (ewaldr_hf.F90:347-371, 438-462),
// and calvirihf (calvirihf.F90:8-405) adds the virial with weight 1/2.
//
// Every pair adds the same kind of term to both of its atoms, and each
// atom's share can be written from that atom's side alone ("target" t, the
// other atom "source" s, force on the target f_t = c (x_t - x_s)):
//   energy:  e_t += E / 2;
//   bulk:    W_t(a, n) += rc_n f_t,a / 2,  rc = minimum image of c_t - c_s;
//   volume:  W_t(a, 3, r) += f_t,a ff(r), ff the signed length of the part of
//            the centres' z segment inside region r (sign of c_t,z - c_s,z)
//            times 1/2;
//   surface: W_t(a, 3, r) += f_t,a sign(c_t,z - z1(r)) / 2 for every plane
//            z1(r) the segment crosses,
// which for t = i is calvirihf's W_i += f_i ..., and for t = j equals its
// W_j += (-f_j) ... with the signs of both factors flipped.  So the pass runs
// twice over the half list, each time adding to the targets only:
//   - by i-cluster (the list as built): the 8 atoms of cluster ci are the
//     targets, the 4 atoms of each listed half-cluster the sources;
//   - by half-cluster over the transposed list (built here from the list, one
//     record per entry, mask bits transposed): the 4 atoms of half-cluster h
//     are the targets, the 8 atoms of each cluster that lists h the sources.
// Targets keep their sums in registers (bulk form; 20 values) or in a per-wave
// LDS slice (volume and surface forms; [target][class][region][3]), and each
// target's sums are added to the device arrays once per pass.  Pairs inside
// one molecule (rare: exclusions remove the bonded ones) add straight to the
// device arrays.
//
// The cutoff decisions are the reference's exact FP64 tests (no contraction,
// r <= rrcut decided as r^2 <= the largest double whose root is <= rrcut); the
// pair terms are FP64, or FP32 in the mixed form (one hardware reciprocal
// root), whose sums are int64 fixed point (order-independent).
#include <hipcub/hipcub.hpp>
#include <type_traits>

#include "pgk_hf.h"

namespace {

struct HfWArgs {
  const int32_t *start;     // by i-cluster: ci_start[ncl+1]; transposed: t_start[2 ncl + 1]
  const int32_t *src;       // by i-cluster: half-cluster index; transposed: cluster index
  const uint32_t *mask;     // bit = lane & 31 of the pass's lane layout
  const double4 *xq;        // per slot: FP64 position, charge
  const double4 *cen;       // per slot: heat-flux centre, w = 1 (ATOM) or 0 (MOLE)
  const int32_t *mol;       // per slot: molecule id (0 for padding)
  const int32_t *ct;        // per slot: 0-based type
  const int32_t *cl_atom;   // per slot: atom index (-1 padding)
  const double *lj;         // ntype^2 x 2: eps (0 unless INTTYPE_VDW), rmin^2
  int nunit, ntype;
  double sqcut, sqcut2, rrsq;
};

__device__ __forceinline__ double r2_exact_d(double x, double y, double z) {
#pragma clang fp contract(off)
  return x * x + y * y + z * z;
}

template <class R> __device__ __forceinline__ R w_rinv(double r2);
template <> __device__ __forceinline__ double w_rinv<double>(double r2) {
  double y = __builtin_amdgcn_rsq(r2);   // seed, then two Newton steps
  const double h = 0.5 * r2;
  y = y * (1.5 - h * y * y);
  y = y * (1.5 - h * y * y);
  return y;
}
template <> __device__ __forceinline__ float w_rinv<float>(double r2) {
  return __builtin_amdgcn_rsqf((float)r2);
}

__device__ __forceinline__ bool w_lane_bit(uint64_t m) {
  return __builtin_amdgcn_inverse_ballot_w64(m);
}

// fixed point or FP64 add into LDS / device memory
__device__ __forceinline__ void w_add(double *p, double v) { atomicAdd(p, v); }
__device__ __forceinline__ void w_add(unsigned long long *p, double v) {
  atomicAdd(p, (unsigned long long)__double2ll_rn(v * PGK_FX_SCALE));
}

// one pair's target-side terms straight into the device arrays (pairs
// inside one molecule), calvirihf restricted to the target
template <class AH>
__device__ void w_direct(const HfDev &h, int at, int cls, double e, const double *f,
                         const double *ct, const double *cs, bool seg) {
  acc_add(reinterpret_cast<AH *>(h.pe) + (size_t)cls * h.natom + at, 0.5 * e);
  if (!seg) return;
  AH *W = reinterpret_cast<AH *>(h.w) + (size_t)cls * h.nreg * h.natom * h.nw;
  const size_t N = (size_t)h.natom;
  if (h.mode == PGK_HF_BULK) {
    double rc[3];
    for (int d = 0; d < 3; d++) {
      rc[d] = ct[d] - cs[d];
      rc[d] = rc[d] - h.L[d] * round(rc[d] * h.Li[d]);
    }
    for (int ir = 0; ir < h.nreg; ir++) {
      AH *wt = W + (ir * N + at) * 9;
      for (int n = 0; n < 3; n++)
        for (int a = 0; a < 3; a++) acc_add(&wt[a + 3 * n], (rc[n] * f[a]) * 0.5);
    }
    return;
  }
  const double r2 = ct[2] - cs[2], rz = fabs(r2);
  if (rz * h.Li[2] >= 0.5) return;
  for (int ir = 0; ir < h.nreg; ir++) {
    double ff;
    if (h.mode == PGK_HF_SURFACE) {
      const double ri1 = ct[2] - h.z1[ir], rj1 = cs[2] - h.z1[ir];
      if (!(ri1 * rj1 < 0.0)) continue;
      ff = ((ri1 > 0.0) ? 1.0 : -1.0) * 0.5;
    } else {
      ff = hf_volume_weight(h, ir, ct[2], cs[2], r2, rz, 0.5);
      if (ff == 0.0) continue;
    }
    AH *wt = W + (ir * N + at) * 3;
    for (int a = 0; a < 3; a++) acc_add(&wt[a], f[a] * ff);
  }
}

// MODE: PGK_HF_BULK, PGK_HF_VOLUME or PGK_HF_SURFACE.  TR: transposed list
// (targets = 4 atoms of a half-cluster, sources = 8 atoms of a cluster).
template <class AH, int MODE, bool TR>
__global__ __launch_bounds__(256) void k_hf_w(HfWArgs p, HfDev h) {
  using R = typename std::conditional<std::is_same<AH, unsigned long long>::value, float,
                                      double>::type;
  constexpr bool BULK = MODE == PGK_HF_BULK;
  constexpr int NT = TR ? 4 : 8;   // targets per wave
  extern __shared__ __align__(16) unsigned char w_smem[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int unit = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wave);
  const int nreg = h.nreg;
  // volume/surface: [NT][2][nreg][3] per wave (inter classes ELC, VDW)
  AH *sl = reinterpret_cast<AH *>(w_smem) + (size_t)wave * NT * 2 * nreg * 3;
  if (!BULK) {   // this wave's own slice: LDS operations of one wave stay in order
    for (int k = lane; k < NT * 2 * nreg * 3; k += 64) sl[k] = AH(0);
  }
  if (unit >= p.nunit) return;
  const int tl = TR ? (lane & 3) : (lane & 7);          // target within the unit
  const int sl_of = TR ? ((lane >> 2) & 7) : ((lane >> 3) & 3);   // source slot within its unit
  const int hs = lane >> 5;                             // entry of the pass
  const int tslot = TR ? 4 * unit + tl : 8 * unit + tl;
  const int at = p.cl_atom[tslot];
  const double4 xt = p.xq[tslot];
  const double4 ctc = p.cen[tslot];
  const int molt = p.mol[tslot];
  const int tyt = p.ct[tslot];
  const double ct3[3] = {ctc.x, ctc.y, ctc.z};
  double pe_c = 0.0, pe_l = 0.0;
  double wc[9], wl[9];
#pragma unroll
  for (int k = 0; k < 9; k++) wc[k] = wl[k] = 0.0;
  const int e0 = __builtin_amdgcn_readfirstlane(p.start[unit]);
  const int e1 = __builtin_amdgcn_readfirstlane(p.start[unit + 1]);
  for (int e2 = e0; e2 < e1; e2 += 2) {
    const uint32_t ma = p.mask[e2];
    const uint32_t mb = (e2 + 1 < e1) ? p.mask[e2 + 1] : 0u;
    const uint64_t m = ((uint64_t)__builtin_amdgcn_readfirstlane((int)mb) << 32) |
                       (uint32_t)__builtin_amdgcn_readfirstlane((int)ma);
    if (!w_lane_bit(m)) continue;
    const int e = e2 + hs;
    const int su = p.src[e];
    const int sslot = TR ? 8 * su + sl_of : 4 * su + sl_of;
    const double4 xs = p.xq[sslot];
    double rx = xt.x - xs.x, ry = xt.y - xs.y, rz = xt.z - xs.z;
    rx = rx - h.L[0] * rint(rx * h.Li[0]);
    ry = ry - h.L[1] * rint(ry * h.Li[1]);
    rz = rz - h.L[2] * rint(rz * h.Li[2]);
    const double r2 = r2_exact_d(rx, ry, rz);
    if (!(r2 <= p.sqcut)) continue;
    const R rinv1 = w_rinv<R>(r2);
    const R rinv2 = rinv1 * rinv1;
    const bool coul = r2 <= p.rrsq;
    R ec = 0, fc = 0, el = 0, fl = 0;
    if (coul) {
      const R qq = (R)xt.w * (R)xs.w;
      ec = qq * rinv1;
      fc = ec * rinv2;
    }
    if (r2 <= p.sqcut2) {
      // type pair in the list's orientation (i type + ntype * j type)
      const int tys = p.ct[sslot];
      const int tp = TR ? tys + p.ntype * tyt : tyt + p.ntype * tys;
      const R eps = (R)p.lj[2 * tp];
      if (eps != R(0)) {
        const R t2 = rinv2 * (R)p.lj[2 * tp + 1];
        const R t6 = t2 * t2 * t2;
        const R t12 = t6 * t6;
        el = eps * (t12 - R(2) * t6);
        fl = R(12) * eps * (t12 - t6) * rinv2;
      }
    }
    const double4 csc = p.cen[sslot];
    const bool same = molt == p.mol[sslot];
    // two MOLE atoms of one molecule add energies but no virial
    const bool seg0 = !(same && ctc.w == 0.0 && csc.w == 0.0);
    const double cs3[3] = {csc.x, csc.y, csc.z};
    if (same) {   // rare: straight to the device arrays
      if (coul) {
        const double f[3] = {(double)fc * rx, (double)fc * ry, (double)fc * rz};
        w_direct<AH>(h, at, PGK_HF_ELIN, (double)ec, f, ct3, cs3, seg0);
      }
      if (el != R(0) || fl != R(0)) {
        const double f[3] = {(double)fl * rx, (double)fl * ry, (double)fl * rz};
        w_direct<AH>(h, at, PGK_HF_LJIN, (double)el, f, ct3, cs3, seg0);
      }
      continue;
    }
    pe_c += 0.5 * (double)ec;
    pe_l += 0.5 * (double)el;
    if (BULK) {
      double rc[3];
      for (int d = 0; d < 3; d++) {
        rc[d] = ct3[d] - cs3[d];
        rc[d] = rc[d] - h.L[d] * round(rc[d] * h.Li[d]);
      }
      const double hc = 0.5 * (double)fc, hl = 0.5 * (double)fl;
      const double fa[3] = {rx, ry, rz};
#pragma unroll
      for (int n = 0; n < 3; n++) {
        const double sc = rc[n] * hc, sv = rc[n] * hl;
#pragma unroll
        for (int a = 0; a < 3; a++) {
          wc[a + 3 * n] += sc * fa[a];
          wl[a + 3 * n] += sv * fa[a];
        }
      }
    } else {
      const double r2z = ct3[2] - cs3[2], az = fabs(r2z);
      if (az * h.Li[2] >= 0.5) continue;
      const double fcx = (double)fc * rx, fcy = (double)fc * ry, fcz = (double)fc * rz;
      const double flx = (double)fl * rx, fly = (double)fl * ry, flz = (double)fl * rz;
      AH *st = sl + (size_t)tl * 2 * nreg * 3;
      const double lo = fmin(ct3[2], cs3[2]), hi = fmax(ct3[2], cs3[2]);
      int r0 = 0, r1 = nreg - 1;
      if (MODE == PGK_HF_VOLUME) hf_region_window(h, lo, hi, r0, r1);
      for (int ir = r0; ir <= r1; ir++) {
        double ff;
        if (MODE == PGK_HF_SURFACE) {
          const double ri1 = ct3[2] - h.z1[ir], rj1 = cs3[2] - h.z1[ir];
          if (!(ri1 * rj1 < 0.0)) continue;
          ff = ((ri1 > 0.0) ? 1.0 : -1.0) * 0.5;
        } else {
          if (hi < h.z1[ir] || lo > h.z2[ir]) continue;   // the segment misses the region
          ff = hf_volume_weight(h, ir, ct3[2], cs3[2], r2z, az, 0.5);
          if (ff == 0.0) continue;
        }
        if (fc != R(0)) {
          w_add(&st[ir * 3 + 0], fcx * ff);
          w_add(&st[ir * 3 + 1], fcy * ff);
          w_add(&st[ir * 3 + 2], fcz * ff);
        }
        if (fl != R(0)) {
          w_add(&st[(nreg + ir) * 3 + 0], flx * ff);
          w_add(&st[(nreg + ir) * 3 + 1], fly * ff);
          w_add(&st[(nreg + ir) * 3 + 2], flz * ff);
        }
      }
    }
  }
  // sums over the lanes that share a target (8 lanes by i-cluster, 16
  // transposed), in a fixed order
  constexpr int G0 = TR ? 4 : 8;   // first lane stride of the group
#pragma unroll
  for (int o = G0; o < 64; o <<= 1) {
    pe_c += __shfl_xor(pe_c, o, 64);
    pe_l += __shfl_xor(pe_l, o, 64);
    if (BULK) {
#pragma unroll
      for (int k = 0; k < 9; k++) {
        wc[k] += __shfl_xor(wc[k], o, 64);
        wl[k] += __shfl_xor(wl[k], o, 64);
      }
    }
  }
  const int gl = lane / G0;               // this lane's index within its target group
  constexpr int GS = 64 / G0;             // lanes per target group
  if (at >= 0) {
    AH *pe = reinterpret_cast<AH *>(h.pe);
    if (gl == 0 && pe_c != 0.0) acc_add(&pe[(size_t)PGK_HF_ELC * h.natom + at], pe_c);
    if (gl == 1 % GS && pe_l != 0.0) acc_add(&pe[(size_t)PGK_HF_VDW * h.natom + at], pe_l);
    if (BULK) {
      AH *Wc = reinterpret_cast<AH *>(h.w) + (size_t)PGK_HF_ELC * nreg * h.natom * 9;
      AH *Wl = reinterpret_cast<AH *>(h.w) + (size_t)PGK_HF_VDW * nreg * h.natom * 9;
      const size_t N = (size_t)h.natom;
#pragma unroll
      for (int k = 0; k < 18; k++) {
        if (k % GS != gl) continue;
        const double v = (k < 9) ? wc[k] : wl[k - 9];
        if (v == 0.0) continue;
        AH *W = (k < 9) ? Wc : Wl;
        for (int ir = 0; ir < nreg; ir++) acc_add(&W[(ir * N + at) * 9 + (k % 9)], v);
      }
    }
  }
  if (!BULK) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    const size_t N = (size_t)h.natom;
    for (int k = lane; k < NT * 2 * nreg * 3; k += 64) {
      const AH v = sl[k];
      if (v == AH(0)) continue;
      const int t = k / (2 * nreg * 3), rem = k - t * 2 * nreg * 3;
      const int q = rem / (nreg * 3), ir = (rem / 3) % nreg, a = rem % 3;
      const int atk = p.cl_atom[TR ? 4 * unit + t : 8 * unit + t];
      if (atk < 0) continue;
      AH *W = reinterpret_cast<AH *>(h.w) +
              (size_t)(q == 0 ? PGK_HF_ELC : PGK_HF_VDW) * nreg * N * 3;
      atomicAdd(&W[(ir * N + atk) * 3 + a], v);
    }
  }
}

}  // namespace

// per slot: FP64 position and charge, heat-flux centre (w = 1 ATOM, 0 MOLE),
// molecule id, and with vel4 the collision-part velocity (h.vel)
__global__ void k_hf_pack(int nslot, const int32_t *__restrict__ cl_atom, HfDev h,
                          const double *__restrict__ q, double4 *__restrict__ xq,
                          double4 *__restrict__ cen, int32_t *__restrict__ mol,
                          double4 *__restrict__ vel4) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslot) return;
  const int a = cl_atom[s];
  if (a < 0) {
    xq[s] = make_double4(0.0, 0.0, 0.0, 0.0);
    cen[s] = make_double4(0.0, 0.0, 0.0, 1.0);
    mol[s] = 0;
    if (vel4) vel4[s] = make_double4(0.0, 0.0, 0.0, 0.0);
    return;
  }
  xq[s] = make_double4(h.pos[3 * a], h.pos[3 * a + 1], h.pos[3 * a + 2], q[a]);
  double c[3];
  const int atomc = hf_centre(h, a, h.mode == PGK_HF_BULK, c);
  cen[s] = make_double4(c[0], c[1], c[2], atomc ? 1.0 : 0.0);
  mol[s] = h.molid[a];
  if (vel4) vel4[s] = make_double4(h.vel[3 * a], h.vel[3 * a + 1], h.vel[3 * a + 2], 0.0);
}

namespace {

// transposed list: entries grouped by half-cluster, each holding its
// cluster and the mask with bit ii + 8 jq moved to bit jq + 4 ii
__device__ __forceinline__ uint32_t transpose_mask(uint32_t m) {
  uint32_t t = 0;
#pragma unroll
  for (int ii = 0; ii < 8; ii++)
#pragma unroll
    for (int jq = 0; jq < 4; jq++) t |= ((m >> (ii + 8 * jq)) & 1u) << (jq + 4 * ii);
  return t;
}

__device__ __forceinline__ int entry_cluster(const int32_t *ci_start, int ncl, int e) {
  int lo = 0, hi = ncl;   // largest ci with ci_start[ci] <= e
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (ci_start[mid] <= e) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ void k_tl_count(int nent, const int32_t *__restrict__ cj, int32_t *__restrict__ cnt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nent) atomicAdd(&cnt[cj[e]], 1);
}

__global__ void k_tl_fill(int nent, int ncl, const int32_t *__restrict__ ci_start,
                          const int32_t *__restrict__ cj, const uint32_t *__restrict__ mask,
                          const int32_t *__restrict__ t_start, int32_t *__restrict__ fill,
                          int32_t *__restrict__ t_src, uint32_t *__restrict__ t_mask) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nent) return;
  const int hc = cj[e];
  const int k = t_start[hc] + atomicAdd(&fill[hc], 1);
  t_src[k] = entry_cluster(ci_start, ncl, e);
  t_mask[k] = transpose_mask(mask[e]);
}

template <class AH, int MODE>
int launch_w_mode(pgk_ctx *c, const HfWArgs &pa, const HfWArgs &pb) {
  hipStream_t st = c->stream;
  const HfDev &h = c->hfdev;
  const size_t lds_a = (MODE == PGK_HF_BULK) ? 0 : (size_t)4 * 8 * 2 * h.nreg * 3 * sizeof(AH);
  const size_t lds_b = (MODE == PGK_HF_BULK) ? 0 : (size_t)4 * 4 * 2 * h.nreg * 3 * sizeof(AH);
  if (lds_a > 0) {
    int lr = pgk_lds_limit((const void *)k_hf_w<AH, MODE, false>, lds_a);
    if (!lr) lr = pgk_lds_limit((const void *)k_hf_w<AH, MODE, true>, lds_b);
    if (lr) return lr;
  }
  k_hf_w<AH, MODE, false><<<pgk_grid_blocks(pa.nunit, 4), 256, lds_a, st>>>(pa, h);
  k_hf_w<AH, MODE, true><<<pgk_grid_blocks(pb.nunit, 4), 256, lds_b, st>>>(pb, h);
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

}  // namespace

// The W-array pair pass of calforce_hf (the context's HfDev in W mode).
// Volume and surface forms keep up to PGK_HF_W_MAXREG regions per wave in
// LDS; more regions are refused by the caller.
int pgk_hf_w_launch(pgk_ctx *c) {
  hipStream_t st = c->stream;
  const HfDev &h = c->hfdev;
  if (c->ncl <= 0) return PGK_OK;
  if (h.mode != PGK_HF_BULK && h.nreg > PGK_HF_W_MAXREG) {
    pgk_set_error("calforce_hf: more than " + std::to_string(PGK_HF_W_MAXREG) +
                  " volume or surface regions");
    return PGK_EINVAL;
  }
  const int ncl = c->ncl, nslot = ncl * PGK_CL, nhc = 2 * ncl;
  const int64_t nent = c->npair_cl;
  int rc;
  if ((rc = grow_dev((void **)&c->d_hfw_xq, &c->hfw_xq_bytes, sizeof(double4) * nslot)) ||
      (rc = grow_dev((void **)&c->d_hfw_cen, &c->hfw_cen_bytes, sizeof(double4) * nslot)) ||
      (rc = grow_dev((void **)&c->d_hfw_mol, &c->hfw_mol_bytes, sizeof(int32_t) * nslot)))
    return rc;
  k_hf_pack<<<pgk_grid_blocks(nslot, 256), 256, 0, st>>>(nslot, c->d_cl_atom, h, c->d_q,
                                                           c->d_hfw_xq, c->d_hfw_cen,
                                                           c->d_hfw_mol, nullptr);
  HIP_CHECK(hipGetLastError());
  if (c->tl_serial != c->list_serial) {   // transposed list of the current list
    if ((rc = grow_dev((void **)&c->d_tl_start, &c->tl_start_bytes,
                       sizeof(int32_t) * (size_t)(2 * nhc + 2))) ||
        (rc = grow_dev((void **)&c->d_tl_src, &c->tl_src_bytes,
                       sizeof(int32_t) * (size_t)(nent > 0 ? nent : 1))) ||
        (rc = grow_dev((void **)&c->d_tl_mask, &c->tl_mask_bytes,
                       sizeof(uint32_t) * (size_t)(nent > 0 ? nent : 1))))
      return rc;
    int32_t *cnt = c->d_tl_start + nhc + 1;   // nhc + 1 counters after the starts
    HIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(int32_t) * (nhc + 1), st));
    if (nent > 0) {
      k_tl_count<<<pgk_grid_blocks(nent, 256), 256, 0, st>>>((int)nent, c->d_cj, cnt);
      HIP_CHECK(hipGetLastError());
    }
    size_t tb = 0;
    HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, c->d_tl_start, nhc + 1, st));
    if ((rc = grow_dev(&c->d_scan_tmp, &c->scan_tmp_bytes, tb))) return rc;
    tb = c->scan_tmp_bytes;
    HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(c->d_scan_tmp, tb, cnt, c->d_tl_start, nhc + 1,
                                               st));
    HIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(int32_t) * (nhc + 1), st));
    if (nent > 0) {
      k_tl_fill<<<pgk_grid_blocks(nent, 256), 256, 0, st>>>(
          (int)nent, ncl, c->d_ci_start, c->d_cj, c->d_mask, c->d_tl_start, cnt, c->d_tl_src,
          c->d_tl_mask);
      HIP_CHECK(hipGetLastError());
    }
    c->tl_serial = c->list_serial;
  }
  HfWArgs pa;
  pa.start = c->d_ci_start;
  pa.src = c->d_cj;
  pa.mask = c->d_mask;
  pa.xq = c->d_hfw_xq;
  pa.cen = c->d_hfw_cen;
  pa.mol = c->d_hfw_mol;
  pa.ct = c->d_ct;
  pa.cl_atom = c->d_cl_atom;
  pa.lj = c->d_ljp64;
  pa.nunit = ncl;
  pa.ntype = c->ntype;
  double sq = c->rrcut * c->rrcut, sq2 = c->rcut * c->rcut;
  pa.sqcut = sq2 > sq ? sq2 : sq;
  pa.sqcut2 = sq2;
  pa.rrsq = c->rrsq_cut;
  HfWArgs pb = pa;
  pb.start = c->d_tl_start;
  pb.src = c->d_tl_src;
  pb.mask = c->d_tl_mask;
  pb.nunit = nhc;
  const bool fx = c->precision == PGK_PREC_MIXED;
  switch (h.mode) {
    case PGK_HF_BULK:
      return fx ? launch_w_mode<unsigned long long, PGK_HF_BULK>(c, pa, pb)
                : launch_w_mode<double, PGK_HF_BULK>(c, pa, pb);
    case PGK_HF_SURFACE:
      return fx ? launch_w_mode<unsigned long long, PGK_HF_SURFACE>(c, pa, pb)
                : launch_w_mode<double, PGK_HF_SURFACE>(c, pa, pb);
    default:
      return fx ? launch_w_mode<unsigned long long, PGK_HF_VOLUME>(c, pa, pb)
                : launch_w_mode<double, PGK_HF_VOLUME>(c, pa, pb);
  }
}
