// k_pair.hip - real-space Ewald + Lennard-Jones pair kernel and the
// excluded-pair correction.
//
// Restates ewaldrp (ewaldr.F90:404-838) for the listed pairs:
//   r^2 <= max(rrcut^2, rcut^2), then
//   Coulomb if r <= rrcut:   f = qq*S_f(r)*r_ij,  E = qq*S_e(r)
//     (S_f, S_e: the reference's natural cubic splines of
//      erfc(ar)/r^3 + 2a/sqrt(pi) e^{-a^2r^2}/r^2 and erfc(ar)/r,
//      spline_interp.F90:96-111, evaluated as SPLINE_INTERPFUNC :196-242)
//   LJ if r^2 <= rcut^2 and inter_inttyp == VDW:
//     E = eps[(R/r)^12 - 2(R/r)^6],  f = 12 eps[(R/r)^12 - (R/r)^6]/r^2 * r_ij
//   virial tensors W(a,n) += f_a r_n per class, scalar LJ virial f.r.
// and the excluded-pair correction (ewaldr.F90:744-828): for every exclusion
//   E -= qq*S'_e(r), f = -qq*S'_f(r)*r_ij with the erf tables, no cutoff.
//
// Mapping: one 64-lane wavefront per 8-atom i-cluster; lane = ii + 8*jj
// handles the pair (i-slot ii, j-slot jj) of the current j-cluster, so a
// cluster-pair mask bit is exactly the lane's predicate.  j-forces are summed
// over the 8 i-lanes with cross-lane shuffles and added once per j-atom; the
// i-forces stay in registers over the whole j list.
#include <type_traits>

#include "pgk_device.h"


template <int PREC>
struct PairArgs {
  using R = typename Prec<PREC>::real;
  using R4 = typename Prec<PREC>::real4;
  using A = typename Prec<PREC>::acc;
  const R4 *xq;
  const int32_t *ct;
  const int32_t *ci_start, *cj;   // entries: half-cluster index 2 cj + h
  const uint32_t *mask;            // entries: bit ii + 8 jq
  const R *spl;          // nspl x 8: force a,b,c,d | energy a,b,c,d
  const R *lj;           // ntype^2 x 2: eps (0 unless INTTYPE_VDW), rmin^2
  A *fcl, *fcl_c;
  double *red;
  const int32_t *witem;  // chunked form: work items (cluster * 32 + chunk)
  int nwi;
  int ncl, ntype, nspl_lds, lj_lds;
  R L[3], Linv[3];
  R sqcut, sqcut2, rrsq;
  R h, hinv;
};

template <class R>
__device__ __forceinline__ void spline2(const R *__restrict__ tb, R x, R h, R hinv, R xmax,
                                        int nspl, R &sf, R &se) {
  if (!(x < xmax)) {
    sf = se = R(0);
    return;
  }
  int i = (int)(x * hinv);
  i = max(0, min(i, nspl - 1));
  R d = x - (R)i * h;
  const R *t = tb + 8 * i;
  sf = t[0] + d * (t[1] + d * (t[2] + d * t[3]));
  se = t[4] + d * (t[5] + d * (t[6] + d * t[7]));
}

// |r|^2 without contraction so the cutoff decision matches the reference
__device__ __forceinline__ double r2_nc(double x, double y, double z) {
#pragma clang fp contract(off)
  return x * x + y * y + z * z;
}
__device__ __forceinline__ float r2_nc(float x, float y, float z) {
#pragma clang fp contract(off)
  return x * x + y * y + z * z;
}

// nearest-image shift; listed pairs are shorter than half the box, where
// round-half-even and the reference's round-half-away agree
__device__ __forceinline__ double nimg(double x) { return rint(x); }
__device__ __forceinline__ float nimg(float x) { return rintf(x); }

// ---- cross-lane sums over the 8 lanes of an aligned group (DPP, no LDS) ----
template <int CTRL>
__device__ __forceinline__ float dppf(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dppd(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
// quad_perm[1,0,3,2] = 0xB1, quad_perm[2,3,0,1] = 0x4E, row_half_mirror = 0x141
__device__ __forceinline__ float sum8(float v) {
  v += dppf<0xB1>(v);
  v += dppf<0x4E>(v);
  v += dppf<0x141>(v);
  return v;
}
__device__ __forceinline__ double sum8(double v) {
  v += dppd<0xB1>(v);
  v += dppd<0x4E>(v);
  v += dppd<0x141>(v);
  return v;
}

// 1/r^2 and r.  FP64: hardware rsq seed and two Newton steps (relative error
// a few ulp; the cutoff decisions never depend on it, see rrsq below).
template <class R>
__device__ __forceinline__ void rinv_pair(R r2, R &rinv2, R &r);
template <>
__device__ __forceinline__ void rinv_pair<double>(double r2, double &rinv2, double &r) {
  double y = __builtin_amdgcn_rsq(r2);
  double h = 0.5 * r2;
  y = y * (1.5 - h * y * y);
  y = y * (1.5 - h * y * y);
  rinv2 = y * y;
  r = r2 * y;
}
template <>
__device__ __forceinline__ void rinv_pair<float>(float r2, float &rinv2, float &r) {
  float ri = __frsqrt_rn(r2);
  rinv2 = ri * ri;
  r = r2 * ri;
}

template <int CTRL>
__device__ __forceinline__ float dpp_t(float v) { return dppf<CTRL>(v); }
template <int CTRL>
__device__ __forceinline__ double dpp_t(double v) { return dppd<CTRL>(v); }

// x, y, z summed over each aligned group of 8 lanes in three exchange steps
// (k = lane & 7): afterwards lanes k = 0|1 hold sum(x), k = 2|3 sum(y) and
// k = 4|5 sum(z).  Step 1 pairs lane k with 7-k (row_half_mirror): the low
// half keeps x and y, the high half z; step 2 pairs k with k^2, step 3 k^1.
template <class R>
__device__ __forceinline__ R sum3x8(R x, R y, R z, int k) {
  const bool hi = (k & 4) != 0;
  const R keep0 = hi ? z : x, send0 = hi ? x : z;
  const R keep1 = hi ? R(0) : y, send1 = hi ? y : R(0);
  const R a0 = keep0 + dpp_t<0x141>(send0);
  const R a1 = keep1 + dpp_t<0x141>(send1);
  const bool b1 = (k & 2) != 0;
  const R b = (b1 ? a1 : a0) + dpp_t<0x4E>(b1 ? a0 : a1);
  return b + dpp_t<0xB1>(b);
}

__device__ __forceinline__ void acc_add_j(unsigned long long *p, float v) {
  atomicAdd(p, fx_of(v));
}
__device__ __forceinline__ void acc_add_j(double *p, double v) { atomicAdd(p, v); }
__device__ __forceinline__ void acc_add_j(float *p, float v) { atomicAdd(p, v); }

// a value every lane holds alike, moved to scalar registers so that the
// arithmetic on it runs on the scalar unit
__device__ __forceinline__ int uni32(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
// bit `lane` of a wave-uniform 64-bit mask, used directly as the lane predicate
__device__ __forceinline__ bool lane_bit(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

// ENE: energies and virial tensors (a call whose energies nobody reads, such
// as an MD step that samples nothing, skips them and writes a zero record).
// CH (chunked form): a wave takes one work item, 64 entries of one cluster's
// list, instead of the whole list, so a small system's few clusters still
// spread over the device; the cluster's i-forces are then added once per
// chunk (fixed point in the mixed form, so the sums do not depend on which
// wave adds them).
template <int PREC, bool SPLIT, bool LJLDS, int NW, bool ENE, bool CH>
__global__ __launch_bounds__(64 * NW) void k_pairs(PairArgs<PREC> a) {
  using R = typename Prec<PREC>::real;
  using R4 = typename Prec<PREC>::real4;
  // per-lane i-force partials: FP64 in the FP64 form; FP32 otherwise (each
  // lane adds its own ~40 pair forces in a fixed order, then the 8 partials
  // of an atom are summed in FP64)
  using IA = typename std::conditional<PREC == PGK_PREC_FP64, double, float>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  R *s_spl = reinterpret_cast<R *>(smem);
  R *s_lj = s_spl + 8 * a.nspl_lds;
  // stage the erfc spline and LJ tables in LDS
  for (int k = threadIdx.x; k < 8 * a.nspl_lds; k += blockDim.x) s_spl[k] = a.spl[k];
  const int nlj = 2 * a.ntype * a.ntype;
  if (LJLDS)
    for (int k = threadIdx.x; k < nlj; k += blockDim.x) s_lj[k] = a.lj[k];
  __syncthreads();
  const R *ljt = LJLDS ? s_lj : a.lj;

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int ci = uni32(blockIdx.x * NW + wave), chunk = -1;
  if (CH) {
    const int w = ci;
    ci = a.ncl;
    if (w < a.nwi) {
      const int item = uni32(a.witem[w]);
      ci = item >> 5;
      chunk = item & 31;
    }
  }
  double ev[14];
#pragma unroll
  for (int k = 0; k < 14; k++) ev[k] = 0.0;
  if (ci < a.ncl) {
    // lane = ii + 8 jq + 32 hs: i-slot ii against j-slot jq of the pass's
    // first (hs = 0) or second (hs = 1) half-cluster entry
    const int ii = lane & 7, jj = lane >> 3, jq = jj & 3, hs = lane >> 5;
    const R4 xi = a.xq[(unsigned)(ci * PGK_CL + ii)];
    const int ti = a.ct[(unsigned)(ci * PGK_CL + ii)];   // 0-based type (0 for padding)
    const int nlast = a.nspl_lds - 1;
    IA fix = 0, fiy = 0, fiz = 0, cix = 0, ciy = 0, ciz = 0;
    R e_elc = 0, e_vdw = 0;
    R tc[6] = {0, 0, 0, 0, 0, 0}, tl[6] = {0, 0, 0, 0, 0, 0};
    int e0 = uni32(a.ci_start[ci]), e1 = uni32(a.ci_start[ci + 1]);
    if (CH) {
      e0 += 64 * chunk;
      e1 = min(e1, e0 + 64);
    }
    // one pass: two half-cluster entries, masks combined as m (bits 0-31 the
    // first, 32-63 the second), hc the lane's half-cluster, xj tj its j-slot
    auto entry = [&](const int hc, const uint64_t m, const R4 xj, const int tj) {
      // every lane runs the same instructions; a lane whose pair is not
      // listed or lies outside a cutoff contributes exact zeros
      R rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
      rx = rx - a.L[0] * nimg(rx * a.Linv[0]);
      ry = ry - a.L[1] * nimg(ry * a.Linv[1]);
      rz = rz - a.L[2] * nimg(rz * a.Linv[2]);
      const R r2 = r2_nc(rx, ry, rz);
      const bool in = lane_bit(m) && (r2 <= a.sqcut);
      const R r2s = in ? r2 : R(1);
      R rinv2, r;
      rinv_pair<R>(r2s, rinv2, r);
      // Coulomb (r <= rrcut: r2 <= rrsq decides it exactly)
      const bool coul = in && (r2 <= a.rrsq);
      int sk = (int)(r * a.hinv);
      sk = min(sk, nlast);
      const R d = r - (R)sk * a.h;
      const R *t = s_spl + 8 * sk;
      const R sf = t[0] + d * (t[1] + d * (t[2] + d * t[3]));
      const R se = t[4] + d * (t[5] + d * (t[6] + d * t[7]));
      const R qq = coul ? xi.w * xj.w : R(0);
      const R ftc = sf * qq;
      if (ENE) e_elc += se * qq;
      // Lennard-Jones (eps is 0 unless the type pair interacts by VDW)
      const bool ljon = in && (r2 <= a.sqcut2);
      const int tp = ti + __mul24(a.ntype, tj);
      const R eps0 = ljt[2 * tp];
      const R eps = ljon ? eps0 : R(0);
      const R rm2 = ljt[2 * tp + 1];
      const R t2 = rinv2 * rm2;
      const R t6 = t2 * t2 * t2;
      const R t12 = t6 * t6;
      const R ftl = R(12) * eps * (t12 - t6) * rinv2;
      if (ENE) {
        e_vdw += eps * (t12 - R(2) * t6);
        const R xx = rx * rx, yy = ry * ry, zz = rz * rz;
        const R xy = rx * ry, xz = rx * rz, yz = ry * rz;
        tc[0] += ftc * xx; tc[1] += ftc * yy; tc[2] += ftc * zz;
        tc[3] += ftc * xy; tc[4] += ftc * xz; tc[5] += ftc * yz;
        tl[0] += ftl * xx; tl[1] += ftl * yy; tl[2] += ftl * zz;
        tl[3] += ftl * xy; tl[4] += ftl * xz; tl[5] += ftl * yz;
      }
      const R ft = ftc + ftl;
      const R fx = ft * rx, fy = ft * ry, fz = ft * rz;
      fix += (IA)fx;
      fiy += (IA)fy;
      fiz += (IA)fz;
      if (SPLIT) {
        cix += (IA)(ftc * rx);
        ciy += (IA)(ftc * ry);
        ciz += (IA)(ftc * rz);
      }
      // j side: x, y, z of each j-slot summed over its 8 i-lanes; lanes
      // ii = 0, 2, 4 add -sum to the x, y, z accumulator of the j atom
      const R g = sum3x8(fx, fy, fz, ii);
      R gc = R(0);
      if (SPLIT) gc = sum3x8(ftc * rx, ftc * ry, ftc * rz, ii);
      // lanes ii = 0, 2, 4 of every j-slot that has a listed pair: bit 8 jj of
      // mb is the OR of byte jj of m
      uint64_t mb = m | (m >> 4);
      mb |= mb >> 2;
      mb |= mb >> 1;
      if (lane_bit((mb & 0x0101010101010101ull) * 0x15ull)) {
        const unsigned s = ((unsigned)hc * 4u + (unsigned)jq) * 3u + (unsigned)(ii >> 1);
        acc_add_j(&a.fcl[s], -g);
        if (SPLIT) acc_add_j(&a.fcl_c[s], -gc);
      }
    };
    // The i-cluster's entries are read 64 at a time into registers (lane k
    // holds entry k of the chunk, the next chunk is already in flight) and
    // handed out with readlane, so a pass needs no memory round trip before
    // its j-slot load is issued.
    struct Ent {
      int hc;
      uint64_t m;
      R4 xj;
      int tj;
    };
    const int nent = e1 - e0;
    int c_hc = (lane < nent) ? a.cj[e0 + lane] : 0;
    uint32_t c_m = (lane < nent) ? a.mask[e0 + lane] : 0u;
    int n_hc = (64 + lane < nent) ? a.cj[e0 + 64 + lane] : 0;
    uint32_t n_m = (64 + lane < nent) ? a.mask[e0 + 64 + lane] : 0u;
    for (int cb = 0; cb < nent; cb += 64) {
      const int nin = min(64, nent - cb);
      // the pass made of chunk entries k and k+1 (mask 0 past the end;
      // the readlane index is clamped, the mask decides)
      auto fetch = [&](Ent &d, int k) {
        const int kk = min(k, 62);
        const int ha = __builtin_amdgcn_readlane(c_hc, kk);
        const int hb = __builtin_amdgcn_readlane(c_hc, kk + 1);
        const uint32_t ma = (k < nin) ? (uint32_t)__builtin_amdgcn_readlane((int)c_m, kk) : 0u;
        const uint32_t mb =
            (k + 1 < nin) ? (uint32_t)__builtin_amdgcn_readlane((int)c_m, kk + 1) : 0u;
        d.m = ((uint64_t)mb << 32) | ma;
        d.hc = hs ? hb : ha;
        const unsigned slot = (unsigned)d.hc * 4u + (unsigned)jq;
        d.xj = a.xq[slot];
        d.tj = a.ct[slot];
      };
      // two passes in flight: each buffer is refilled right after its pass
      // is computed, while the other buffer's pass is computed.  Without
      // energies the second pass is not guarded by k + 2 < nin: past the
      // chunk's end its mask is 0, so it adds exact zeros and no j-atomics,
      // and the loop body stays one straight path.  With a guard, the
      // wait-count analysis merges the two paths and makes the first pass
      // wait for the second buffer's loads (and the previous pass's
      // j-atomics) as well as its own: the MD step's kernel takes 10 % longer.
      // The energy kernels keep the guard: unguarded, their interleaved
      // passes need more registers and run at lower occupancy.
      Ent A, B;
      fetch(A, 0);
      fetch(B, 2);
      for (int k = 0; k < nin; k += 4) {
        entry(A.hc, A.m, A.xj, A.tj);
        fetch(A, k + 4);
        if (!ENE || k + 2 < nin) {
          entry(B.hc, B.m, B.xj, B.tj);
          fetch(B, k + 6);
        }
      }
      c_hc = n_hc;
      c_m = n_m;
      n_hc = (cb + 128 + lane < nent) ? a.cj[e0 + cb + 128 + lane] : 0;
      n_m = (cb + 128 + lane < nent) ? a.mask[e0 + cb + 128 + lane] : 0u;
    }
    // i side: sum over the 8 j-lanes that share this i-slot (FP64)
    double gx = fix, gy = fiy, gz = fiz, hx = cix, hy = ciy, hz = ciz;
#pragma unroll
    for (int o = 8; o < 64; o <<= 1) {
      gx += __shfl_xor(gx, o, 64);
      gy += __shfl_xor(gy, o, 64);
      gz += __shfl_xor(gz, o, 64);
      if (SPLIT) {
        hx += __shfl_xor(hx, o, 64);
        hy += __shfl_xor(hy, o, 64);
        hz += __shfl_xor(hz, o, 64);
      }
    }
    if (jj < 3) {   // lanes jj = 0, 1, 2 hold x, y, z of slot ii
      const unsigned s = (unsigned)(ci * PGK_CL + ii) * 3u;
      const double v = (jj == 0) ? gx : ((jj == 1) ? gy : gz);
      acc_add(&a.fcl[s + jj], v);
      if (SPLIT) {
        const double w = (jj == 0) ? hx : ((jj == 1) ? hy : hz);
        acc_add(&a.fcl_c[s + jj], w);
      }
    }
    if (ENE) {
      ev[0] = e_elc;
      ev[1] = e_vdw;
#pragma unroll
      for (int k = 0; k < 6; k++) {
        ev[2 + k] = tc[k];
        ev[8 + k] = tl[k];
      }
    }
  }
  if (ENE)
    block_record<14, NW>(ev, a.red + (size_t)blockIdx.x * PGK_REC);
  else if (threadIdx.x < 14)
    a.red[(size_t)blockIdx.x * PGK_REC + threadIdx.x] = 0.0;
}

// ---- excluded pairs, then the force combination -----------------------------
// k_excl: one thread per atom, the excluded-pair correction
// (ewaldr.F90:744-828: E -= qq S'_e(r), f = -qq S'_f(r) r_ij with the erf
// tables, no cutoff) over the atom's symmetric exclusion list, added to the
// atom-order accumulators (and the Coulomb class's).  Each pair is evaluated
// in the reference's orientation (lower index first), so both atoms see the
// same value; the lower atom records the pair's energy and virial.  FP64
// arithmetic in every form.  It reads only positions and charges, so it can
// run beside the pair kernel.
// k_combine: one thread per atom, the slot-order pair forces and the
// atom-order accumulators (PME, bonded, excluded pairs) summed into the FP64
// output and cleared for the next call.  No atomics.
struct CombineArgs {
  int n, use_cl;
  const int32_t *slot_of;
  const int32_t *xs_start, *xs_j;   // symmetric exclusion CSR, 0-based
  const double *pos, *q, *splx;
  double h, hinv, xmax;
  int nspl;
  double3 L, Li;
  double *red;
};

template <class A, bool SPLIT>
__global__ __launch_bounds__(256) void k_excl(CombineArgs a, A *__restrict__ fat,
                                              A *__restrict__ fatc) {
  double ev[7] = {0, 0, 0, 0, 0, 0, 0};
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.n && a.xs_start[i] < a.xs_start[i + 1]) {
    double fe[3] = {0.0, 0.0, 0.0};
    for (int k = a.xs_start[i]; k < a.xs_start[i + 1]; k++) {
      const int j = a.xs_j[k];
      const int lo = min(i, j), hi = max(i, j);
      double rx = a.pos[3 * lo] - a.pos[3 * hi], ry = a.pos[3 * lo + 1] - a.pos[3 * hi + 1],
             rz = a.pos[3 * lo + 2] - a.pos[3 * hi + 2];
      rx = rx - a.L.x * fnint(rx * a.Li.x);
      ry = ry - a.L.y * fnint(ry * a.Li.y);
      rz = rz - a.L.z * fnint(rz * a.Li.z);
      const double r = sqrt(r2_nc(rx, ry, rz));
      double sf, se;
      spline2(a.splx, r, a.h, a.hinv, a.xmax, a.nspl, sf, se);
      const double qq = a.q[lo] * a.q[hi];
      const double ft = -sf * qq;
      const double fx = ft * rx, fy = ft * ry, fz = ft * rz;
      if (i == lo) {
        fe[0] += fx;
        fe[1] += fy;
        fe[2] += fz;
        ev[0] -= se * qq;
        ev[1] += fx * rx; ev[2] += fy * ry; ev[3] += fz * rz;
        ev[4] += fx * ry; ev[5] += fx * rz; ev[6] += fy * rz;
      } else {
        fe[0] -= fx;
        fe[1] -= fy;
        fe[2] -= fz;
      }
    }
    for (int d = 0; d < 3; d++) {
      acc_add(&fat[3 * i + d], fe[d]);
      if (SPLIT) acc_add(&fatc[3 * i + d], fe[d]);
    }
  }
  block_record<7, 4>(ev, a.red + (size_t)blockIdx.x * PGK_REC);
}

template <class A, bool SPLIT>
__global__ __launch_bounds__(256) void k_combine(CombineArgs a, A *__restrict__ fcl,
                                                 A *__restrict__ fat, A *__restrict__ fclc,
                                                 A *__restrict__ fatc, double *__restrict__ force,
                                                 double *__restrict__ fvc,
                                                 double *__restrict__ fvl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int s = a.use_cl ? a.slot_of[i] : -1;
  // every accumulator is read once here and left zeroed for the next call
  for (int d = 0; d < 3; d++) {
    double fp = 0.0;
    if (s >= 0) {
      fp = acc_load(&fcl[3 * s + d]);
      fcl[3 * s + d] = A(0);
    }
    force[3 * i + d] = fp + acc_load(&fat[3 * i + d]);
    fat[3 * i + d] = A(0);
    if (SPLIT) {
      double fc = 0.0;
      if (s >= 0) {
        fc = acc_load(&fclc[3 * s + d]);
        fclc[3 * s + d] = A(0);
      }
      fvc[3 * i + d] = fc + acc_load(&fatc[3 * i + d]);
      fatc[3 * i + d] = A(0);
      fvl[3 * i + d] = fp - fc;
    }
  }
}

template <int PREC>
static int launch_pairs(pgk_ctx *c, bool split, bool ene, int *off, int *nrec) {
  using R = typename Prec<PREC>::real;
  using R4 = typename Prec<PREC>::real4;
  using A = typename Prec<PREC>::acc;
  constexpr int NW = 8;
  PairArgs<PREC> a;
  a.xq = reinterpret_cast<const R4 *>(c->d_xq);
  a.ct = c->d_ct;
  a.ci_start = c->d_ci_start;
  a.cj = c->d_cj;
  a.mask = c->d_mask;
  if (PREC == PGK_PREC_FP64) {
    a.spl = reinterpret_cast<const R *>(c->d_spl);
    a.lj = reinterpret_cast<const R *>(c->d_ljp64);
  } else {
    a.spl = reinterpret_cast<const R *>(c->d_splf);
    a.lj = reinterpret_cast<const R *>(c->d_ljp32);
  }
  a.fcl = reinterpret_cast<A *>(c->d_fcl);
  a.fcl_c = reinterpret_cast<A *>(c->d_fcl_c);
  a.witem = c->d_witem;
  a.nwi = c->nwi;
  a.ncl = c->ncl;
  a.ntype = c->ntype;
  // only the table rows up to rrcut are ever read
  int nused = (int)(c->rrcut / c->spl_int) + 2;
  a.nspl_lds = nused < c->nspl ? nused : c->nspl;
  size_t lds = (size_t)8 * a.nspl_lds * sizeof(R);
  size_t ljb = (size_t)2 * c->ntype * c->ntype * sizeof(R);
  a.lj_lds = (lds + ljb <= 96 * 1024) ? 1 : 0;
  if (a.lj_lds) lds += ljb;
  // beside the reciprocal chain, the pair blocks leave room on every CU: an
  // LDS request above 1/(B+1) of the CU's LDS caps them at B blocks (of NW
  // waves) per CU, so the chain's short kernels find free slots at once
  // (100k water box, mixed MD step: 0.291 ms at B = 2, 0.325 ms at B = 3, 4,
  // 5 or no cap; the FP64 step is the same at every B)
  if (c->pair_cap_blocks > 0 && c->lds_per_cu > 0)
    lds = std::max(lds, (size_t)c->lds_per_cu / (size_t)(c->pair_cap_blocks + 1) + 1024);
  for (int d = 0; d < 3; d++) {
    a.L[d] = (R)c->box[d];
    a.Linv[d] = (R)(1.0 / c->box[d]);
  }
  double sq = c->rrcut * c->rrcut, sq2 = c->rcut * c->rcut;
  if (sq2 > sq) sq = sq2;
  a.sqcut = (R)sq;
  a.sqcut2 = (R)sq2;
  a.rrsq = (R)c->rrsq_cut;
  a.h = (R)c->spl_int;
  a.hinv = (R)(1.0 / c->spl_int);
  // chunked form for small systems, whose clusters do not fill the device
  // with one wave each
  const bool ch = pgk_pair_chunked(c);
  int nblk = pgk_grid_blocks(pgk_pair_waves(c), NW);
  *off = c->red_used;
  *nrec = nblk;
  a.red = c->d_red + (size_t)c->red_used * PGK_REC;
  c->red_used += nblk;
#define PGK_LAUNCH_PAIRS_C(SP, LL, EN, CHK)                                             \
  {                                                                                      \
    auto kf = k_pairs<PREC, SP, LL, NW, EN, CHK>;                                        \
    {                                                                                    \
      const int lr = pgk_lds_limit((const void *)kf, lds);                               \
      if (lr) return lr;                                                                 \
    }                                                                                    \
    kf<<<nblk, 64 * NW, lds, c->stream>>>(a);                                            \
  }
#define PGK_LAUNCH_PAIRS(SP, LL, EN)                                                    \
  {                                                                                      \
    if (ch) PGK_LAUNCH_PAIRS_C(SP, LL, EN, true) else PGK_LAUNCH_PAIRS_C(SP, LL, EN, false) \
  }
  if (split) {   // the per-class forces come with their energies
    if (a.lj_lds) PGK_LAUNCH_PAIRS(true, true, true) else PGK_LAUNCH_PAIRS(true, false, true)
  } else if (ene) {
    if (a.lj_lds) PGK_LAUNCH_PAIRS(false, true, true) else PGK_LAUNCH_PAIRS(false, false, true)
  } else {
    if (a.lj_lds) PGK_LAUNCH_PAIRS(false, true, false) else PGK_LAUNCH_PAIRS(false, false, false)
  }
#undef PGK_LAUNCH_PAIRS
#undef PGK_LAUNCH_PAIRS_C
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

int pgk_pairs_launch(pgk_ctx *c, bool split, bool ene, int *off, int *nrec) {
  switch (c->precision) {
    case PGK_PREC_FP64: return launch_pairs<PGK_PREC_FP64>(c, split, ene, off, nrec);
    case PGK_PREC_MIXED: return launch_pairs<PGK_PREC_MIXED>(c, split, ene, off, nrec);
    default: return launch_pairs<PGK_PREC_FP32>(c, split, ene, off, nrec);
  }
}

static CombineArgs combine_args(pgk_ctx *c, bool do_pair) {
  CombineArgs a;
  a.n = c->natom;
  a.use_cl = do_pair ? 1 : 0;
  a.slot_of = c->d_slot_of;
  a.xs_start = c->d_xs_start;
  a.xs_j = c->d_xs_j;
  a.pos = c->d_pos;
  a.q = c->d_q;
  a.splx = c->d_splx;
  a.h = c->spl_int;
  a.hinv = 1.0 / c->spl_int;
  a.xmax = c->spl_int * (double)(c->nspl - 1);
  a.nspl = c->nspl;
  a.L = make_double3(c->box[0], c->box[1], c->box[2]);
  a.Li = make_double3(1.0 / c->box[0], 1.0 / c->box[1], 1.0 / c->box[2]);
  a.red = nullptr;
  return a;
}

template <class A>
static int launch_excl(pgk_ctx *c, bool split, int *off, int *nrec) {
  CombineArgs a = combine_args(c, true);
  const int nblk = pgk_grid_blocks(c->natom, 256);
  *off = c->red_used;
  *nrec = nblk;
  a.red = c->d_red + (size_t)c->red_used * PGK_REC;
  c->red_used += nblk;
  A *fat = reinterpret_cast<A *>(c->d_fat), *fatc = reinterpret_cast<A *>(c->d_fat_c);
  if (split)
    k_excl<A, true><<<nblk, 256, 0, c->stream>>>(a, fat, fatc);
  else
    k_excl<A, false><<<nblk, 256, 0, c->stream>>>(a, fat, fatc);
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

// The excluded-pair correction of the listed pair terms into the atom-order
// accumulators, with its energy/virial records (*nrec = 0 when the system
// has no exclusions).
int pgk_excl_launch(pgk_ctx *c, bool split, int *off, int *nrec) {
  *nrec = 0;
  if (c->nexclpair <= 0) return PGK_OK;
  switch (c->precision) {
    case PGK_PREC_FP64: return launch_excl<double>(c, split, off, nrec);
    case PGK_PREC_MIXED: return launch_excl<unsigned long long>(c, split, off, nrec);
    default: return launch_excl<float>(c, split, off, nrec);
  }
}

template <class A>
static int launch_combine(pgk_ctx *c, bool do_pair, bool split, double *fout, double *fvc,
                          double *fvl) {
  const CombineArgs a = combine_args(c, do_pair);
  const int nblk = pgk_grid_blocks(c->natom, 256);
  A *fcl = reinterpret_cast<A *>(c->d_fcl), *fat = reinterpret_cast<A *>(c->d_fat);
  A *fclc = reinterpret_cast<A *>(c->d_fcl_c), *fatc = reinterpret_cast<A *>(c->d_fat_c);
  if (split)
    k_combine<A, true><<<nblk, 256, 0, c->stream>>>(a, fcl, fat, fclc, fatc, fout, fvc, fvl);
  else
    k_combine<A, false><<<nblk, 256, 0, c->stream>>>(a, fcl, fat, fclc, fatc, fout, fvc, fvl);
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

// Sum every force accumulator into fout (and the per-class forces when
// split).
int pgk_combine_launch(pgk_ctx *c, bool do_pair, bool split, double *fout, double *fvc,
                       double *fvl) {
  switch (c->precision) {
    case PGK_PREC_FP64: return launch_combine<double>(c, do_pair, split, fout, fvc, fvl);
    case PGK_PREC_MIXED:
      return launch_combine<unsigned long long>(c, do_pair, split, fout, fvc, fvl);
    default: return launch_combine<float>(c, do_pair, split, fout, fvc, fvl);
  }
}
