// k_hf.hip - fused heat-flux pass over the listed pairs (pgk_calforce_heatf
// and the MD loop's samples), and the export of calforce_hf's per-atom arrays
// (whose pair pass is k_hf_w.hip).
//
// Restates the per-atom part of ewaldrp_hf (ewaldr_hf.F90:8-670) in the
// default build: for every listed pair with r <= rrcut the bare Coulomb
// energy qq/r and force qq/r^3 r_ij (ewaldr_hf.F90:336-343), for r^2 <= rcut^2
// and INTTYPE_VDW the LJ energy and force; half of each energy goes to each
// atom, class ELIN/LJIN for pairs inside one molecule and ELC/VDW otherwise
// (ewaldr_hf.F90:347-371, 438-462), and calvirihf scatters the virial with
// weight 1/2.  Excluded pairs and the reciprocal sum add no per-atom terms in
// this build (ewaldr_hf.F90:637-653 only under _HF_BULK_EWK).
//
// The cutoff tests are the reference's in every precision form (FP64, no
// contraction, r = sqrt(r^2), r <= rrcut); the pair terms are FP64, or FP32
// in the mixed form.  The pair terms
// are symmetric under i <-> j, so the cluster list's orientation of a pair
// does not matter.  One 64-lane wave per i-cluster (or per 64 entries of one
// in the chunked form), lane = ii + 8 jj.
#include <type_traits>

#include "pgk_hf.h"

namespace {

struct HfPairArgs {
  const int32_t *ci_start, *cj, *cl_atom;   // cj: half-cluster index of each entry
  const uint32_t *mask;
  // per cluster slot (k_hf_pack): position and charge, centre (w = 1 ATOM,
  // 0 MOLE), collision-part velocity, molecule id, 0-based type
  const double4 *xq, *cen, *v4;
  const int32_t *mol, *ct;
  const double *lj;       // ntype^2 x (eps, 0 unless INTTYPE_VDW; rmin^2)
  const int32_t *witem;   // chunked form: work items (cluster * 32 + chunk of 64 entries)
  int ncl, nwi, ntype;
  double sqcut, sqcut2, rrcut;
  double rrsq;   // largest r^2 with sqrt(r^2) <= rrcut: r2 <= rrsq decides r <= rrcut exactly
};

__device__ __forceinline__ double r2_exact(const double *r) {
#pragma clang fp contract(off)
  return r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
}

// Flux form of the pass (pgk_calforce_heatf): the same pairs, energies and
// per-atom energy sums, but the collision part of calheatf is summed on the
// spot, s = fi . (v_i + v_j) times the region weights, instead of scattering
// W arrays.  Bulk: 4 pair classes x 3 components per lane in registers.
// Volume: a [4][nreg] slice of LDS per wave, added to with LDS atomics.  Each
// block then adds its waves in order into jpart[block].
constexpr int kPairCls0 = PGK_HF_ELIN;   // ELIN, LJIN, ELC, VDW are 4..7

// what an interaction needs besides its class, energy and force coefficient
struct FluxPair {
  bool seg;          // the centres' segment counts (not two MOLE atoms of one molecule...)
  double rw;         // r . (v_i + v_j)
  double w[3];       // v_i + v_j
  double rc[3];      // centre separation (minimum image in the bulk form)
  double ciz, cjz, rz;
};

// one interaction of pair class Q (0 ELIN, 1 LJIN, 2 ELC, 3 VDW) with energy
// e and force coefficient f (force on i = f r): half the energy to each
// atom, and f (r . (v_i + v_j)) / 2 times the region weights to the flux.
// DT (bulk form over the whole box, atom centres, transport velocities =
// the pair's velocities): the energies' share of the transport part,
// e/2 (v_i + v_j), is summed here (tr) instead of per atom.
template <int Q, bool BULK, bool DT>
__device__ __forceinline__ void flux_add(const FluxPair &fp, double e, double f,
                                         double (&ei)[4], double (&ej)[4],
                                         double (&acc)[4][3], double (&tr)[4][3],
                                         double &sv_out) {
  if (DT) {
    const double he = 0.5 * e;
#pragma unroll
    for (int n = 0; n < 3; n++) tr[Q][n] += he * fp.w[n];
  } else {
    ei[Q] += 0.5 * e;
    ej[Q] += 0.5 * e;
  }
  if (!fp.seg) return;
  const double sv = f * fp.rw * 0.5;
  if (BULK) {
#pragma unroll
    for (int n = 0; n < 3; n++) acc[Q][n] += fp.rc[n] * sv;
  } else {
    sv_out = sv;   // the region walk adds it once per pair, for both classes
  }
}

// volume form: the pair's Coulomb (class qc) and LJ (class ql) shares times
// the weight of every region its centres' segment crosses (zr: the regions'
// bounds, z1 then z2, staged in LDS)
__device__ __forceinline__ void flux_regions(const HfDev &h, const FluxPair &fp, int qc,
                                             double svc, int ql, double svl, double *sl,
                                             const double *zr) {
  if (!fp.seg || (svc == 0.0 && svl == 0.0)) return;
  const double lo = fmin(fp.ciz, fp.cjz), hi = fmax(fp.ciz, fp.cjz);
  if (h.tiled) {
    // ka, kb: the last region with z1 <= lo, hi (-1: none).  Regions below
    // ka - 1 end below lo, regions above kb start above hi: the weight
    // misses them.  ka - 1 ends at z1[ka]: it is weighed only when that is
    // lo itself.  Regions strictly between ka and kb have lo < z1 < z2 <= hi,
    // so both planes are crossed (or the upper plane holds the upper centre)
    // and the weight is their signed width, as hf_volume_weight_z gives it.
    const int nreg = h.nreg;
    int ka = min(nreg - 1, max(-1, (int)floor((lo - h.ubz0) * h.ubwi)));
    while (ka + 1 < nreg && zr[ka + 1] <= lo) ka++;
    while (ka >= 0 && zr[ka] > lo) ka--;
    int kb = min(nreg - 1, max(ka, (int)floor((hi - h.ubz0) * h.ubwi)));
    while (kb + 1 < nreg && zr[kb + 1] <= hi) kb++;
    while (kb > ka && zr[kb] > hi) kb--;
    auto exact = [&](int ir) {
      const double z1 = zr[ir], z2 = zr[nreg + ir];
      if (hi < z1 || lo > z2) return;
      const double ff = hf_volume_weight_z(z1, z2, fp.ciz, fp.cjz, fp.rc[2], fp.rz, 1.0);
      if (ff == 0.0) return;
      if (svc != 0.0) atomicAdd(&sl[qc * nreg + ir], ff * svc);
      if (svl != 0.0) atomicAdd(&sl[ql * nreg + ir], ff * svl);
    };
    if (ka >= 1 && zr[ka] == lo) exact(ka - 1);
    if (ka >= 0) exact(ka);
    const bool neg = !(fp.rc[2] >= 0.0);
    for (int ir = max(ka + 1, 0); ir < kb; ir++) {
      const double wd = zr[nreg + ir] - zr[ir];
      const double ff = neg ? -wd : wd;
      if (svc != 0.0) atomicAdd(&sl[qc * nreg + ir], ff * svc);
      if (svl != 0.0) atomicAdd(&sl[ql * nreg + ir], ff * svl);
    }
    if (kb > ka) exact(kb);
    return;
  }
  int r0, r1;
  hf_region_window(h, lo, hi, r0, r1);
  for (int ir = r0; ir <= r1; ir++) {
    const double z1 = zr[ir], z2 = zr[h.nreg + ir];
    if (hi < z1 || lo > z2) continue;   // the segment misses the region
    const double ff = hf_volume_weight_z(z1, z2, fp.ciz, fp.cjz, fp.rc[2], fp.rz, 1.0);
    if (ff == 0.0) continue;
    if (svc != 0.0) atomicAdd(&sl[qc * h.nreg + ir], ff * svc);
    if (svl != 0.0) atomicAdd(&sl[ql * h.nreg + ir], ff * svl);
  }
}

// Cross-lane exchange inside aligned groups of 8 lanes (DPP, no LDS):
// quad_perm [1,0,3,2] = 0xB1, quad_perm [2,3,0,1] = 0x4E, row_half_mirror
// (lane k <-> 7 - k) = 0x141.
template <int CTRL>
__device__ __forceinline__ double dpp8d(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// e[0..3] summed over each aligned group of 8 lanes (k = lane & 7) in four
// exchanges: afterwards lanes k and k^1 hold the sum of e[k >> 1].  Step 1
// pairs k with 7 - k (the low half keeps e0 and e1, the high half e2 and
// e3), step 2 pairs k with k^2 (k & 2 keeps the second of its two), step 3
// k with k^1.  The group's whole wave must be active.
__device__ __forceinline__ double sum4x8(const double (&e)[4], int k) {
  const bool hi = (k & 4) != 0;
  const double a0 = (hi ? e[2] : e[0]) + dpp8d<0x141>(hi ? e[0] : e[2]);
  const double a1 = (hi ? e[3] : e[1]) + dpp8d<0x141>(hi ? e[1] : e[3]);
  const bool b1 = (k & 2) != 0;
  const double b = (b1 ? a1 : a0) + dpp8d<0x4E>(b1 ? a0 : a1);
  return b + dpp8d<0xB1>(b);
}

// 1/sqrt(r^2) in the compute type: FP64 divide of the root, or FP32's
// hardware reciprocal root (about one ulp)
template <class R> __device__ __forceinline__ R hf_rinv(double r2);
template <> __device__ __forceinline__ double hf_rinv<double>(double r2) { return 1.0 / sqrt(r2); }
template <> __device__ __forceinline__ float hf_rinv<float>(double r2) {
  return __builtin_amdgcn_rsqf((float)r2);
}

// CH: the chunked form, one wave per work item (64 list entries of one
// cluster) instead of one wave per cluster; the i atom's energies are then
// added once per item
template <class AH, bool BULK, bool CH, bool DT>
__global__ __launch_bounds__(256) void k_hf_pairs_flux(HfPairArgs p, HfDev h) {
  // the mixed form (fixed-point accumulators) computes the pair terms in FP32
  using R = typename std::conditional<std::is_same<AH, unsigned long long>::value, float,
                                      double>::type;
  extern __shared__ double lds[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int ci = blockIdx.x * 4 + wave, chunk = 0;
  if (CH) {
    const int w = ci;
    ci = p.ncl;
    if (w < p.nwi) {
      const int item = p.witem[w];
      ci = item >> 5;
      chunk = item & 31;
    }
  }
  const int nreg = h.nreg;
  double acc[4][3] = {}, tr[4][3] = {};
  double *sl = lds + (size_t)wave * 4 * nreg;   // volume: this wave's slice
  double *zr = lds + (size_t)16 * nreg;           // volume: the regions' bounds
  if (!BULK) {
    for (int k = lane; k < 4 * nreg; k += 64) sl[k] = 0.0;
    for (int k = threadIdx.x; k < nreg; k += 256) {
      zr[k] = h.z1[k];
      zr[nreg + k] = h.z2[k];
    }
    __syncthreads();
  }
  if (ci < p.ncl) {
    const int ii = lane & 7, jq = (lane >> 3) & 3, hs = lane >> 5;
    // the i atom's slot, loaded once for the whole list (padding slots, ai =
    // -1, have no mask bits)
    const int islot = ci * PGK_CL + ii;
    const int ai = p.cl_atom[islot];
    const double4 xi = p.xq[islot], cci = p.cen[islot], vi = p.v4[islot];
    const int moli = p.mol[islot], typi = p.ct[islot];
    // per-atom energies: the i atom's shares stay in registers for the whole
    // list; the j atom's are summed over the 8 lanes that share it (lanes
    // ii = 0..7 of one (jq, hs) group) and added once per entry
    double ei[4] = {0.0, 0.0, 0.0, 0.0};   // ELIN, LJIN, ELC, VDW
    int e0 = p.ci_start[ci], e1 = p.ci_start[ci + 1];
    if (CH) {
      e0 += 64 * chunk;
      e1 = min(e1, e0 + 64);
    }
    for (int e2 = e0; e2 < e1; e2 += 2) {
      const int e = e2 + hs;
      const bool ein = e < e1;
      const uint32_t m = ein ? p.mask[e] : 0u;
      const int js = ein ? p.cj[e] * 4 + jq : 0;
      const int aj = ein ? p.cl_atom[js] : -1;
      double ej[4] = {0.0, 0.0, 0.0, 0.0};
      if ((m >> (lane & 31)) & 1u) {
        const double4 xj = p.xq[js];
        double r[3] = {xi.x - xj.x, xi.y - xj.y, xi.z - xj.z};
        for (int d = 0; d < 3; d++) r[d] = r[d] - h.L[d] * rint(r[d] * h.Li[d]);
        const double r2 = r2_exact(r);
        if (r2 <= p.sqcut) {
          // the cutoff tests stay exact FP64 decisions; the pair terms are in
          // the compute type R (FP32 in the mixed form: one hardware rsqrt)
          const R rinv1 = hf_rinv<R>(r2);
          const R rinv2 = rinv1 * rinv1;
          const bool same = moli == p.mol[js];
          const double4 ccj = p.cen[js];
          // the centres' separation and, in the volume form, whether it counts
          double rc[3], rz = 0.0;
          bool seg = true;
          if (BULK && cci.w != 0.0 && ccj.w != 0.0) {
            // two atom centres in the bulk form: the centres' minimum image
            // is the pair's own r (round and rint differ only at a
            // half-box tie, which no listed pair reaches)
            for (int d = 0; d < 3; d++) rc[d] = r[d];
          } else {
            seg = !(cci.w == 0.0 && ccj.w == 0.0 && same);
            rc[0] = cci.x - ccj.x;
            rc[1] = cci.y - ccj.y;
            rc[2] = cci.z - ccj.z;
            if (BULK) {
              for (int d = 0; d < 3; d++) rc[d] = rc[d] - h.L[d] * round(rc[d] * h.Li[d]);
            } else {
              rz = fabs(rc[2]);
              seg = seg && rz * h.Li[2] < 0.5;
            }
          }
          // s = fi . v_i - fj . v_j with fj = -fi, per unit force coefficient
          const double4 vj = p.v4[js];
          const double w[3] = {vi.x + vj.x, vi.y + vj.y, vi.z + vj.z};
          const double rw = r[0] * w[0] + r[1] * w[1] + r[2] * w[2];
          const FluxPair fp{seg, rw, {w[0], w[1], w[2]}, {rc[0], rc[1], rc[2]}, cci.z, ccj.z, rz};
          double svc = 0.0, svl = 0.0;
          if (r2 <= p.rrsq) {
            const R qq = (R)xi.w * (R)xj.w;
            const R ph = qq * rinv1;
            if (same) flux_add<0, BULK, DT>(fp, ph, ph * rinv2, ei, ej, acc, tr, svc);
            else flux_add<2, BULK, DT>(fp, ph, ph * rinv2, ei, ej, acc, tr, svc);
          }
          if (r2 <= p.sqcut2) {
            const int tp = typi + p.ntype * p.ct[js];
            const R ep = (R)p.lj[2 * tp];
            if (ep != R(0)) {   // INTTYPE_VDW pairs only
              const R t2 = rinv2 * (R)p.lj[2 * tp + 1];
              const R t6 = t2 * t2 * t2;
              const R t12 = t6 * t6;
              const R pot = ep * (t12 - (R)2 * t6);
              const R ft = (R)12 * ep * (t12 - t6) * rinv2;
              if (same) flux_add<1, BULK, DT>(fp, pot, ft, ei, ej, acc, tr, svl);
              else flux_add<3, BULK, DT>(fp, pot, ft, ei, ej, acc, tr, svl);
            }
          }
          if (!BULK) flux_regions(h, fp, same ? 0 : 2, svc, same ? 1 : 3, svl, sl, zr);
        }
      }
      if (!DT) {
        // the j atom's four class energies summed over its 8 lanes; lanes
        // ii and ii^1 then hold class (ii >> 1) and one of them adds it
        const double sj4 = sum4x8(ej, ii);
        if (aj >= 0 && (ii & 1) == 0 && sj4 != 0.0)
          hf_pe_add<AH>(h, kPairCls0 + (ii >> 1), aj, sj4);
      }
    }
    if (!DT && ai >= 0) {
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (ei[q] != 0.0) hf_pe_add<AH>(h, kPairCls0 + q, ai, ei[q]);
    }
  }
  if (BULK) {
    // per block: the 12 flux sums, then (DT) the 12 transport sums, else 0
    __shared__ double wsum[4][24];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int n = 0; n < 3; n++) {
        const double t = wave_sum(acc[q][n]);
        const double u = DT ? wave_sum(tr[q][n]) : 0.0;
        if (lane == 0) {
          wsum[wave][3 * q + n] = t;
          wsum[wave][12 + 3 * q + n] = u;
        }
      }
    __syncthreads();
    if (threadIdx.x < 24)
      h.jpart[(size_t)blockIdx.x * 24 + threadIdx.x] =
          ((wsum[0][threadIdx.x] + wsum[1][threadIdx.x]) + wsum[2][threadIdx.x]) +
          wsum[3][threadIdx.x];
  } else {
    __syncthreads();
    const int nv = 4 * nreg;
    for (int k = threadIdx.x; k < nv; k += 256)
      h.jpart[(size_t)blockIdx.x * nv + k] =
          ((lds[k] + lds[nv + k]) + lds[2 * nv + k]) + lds[3 * nv + k];
  }
}

// accumulators -> doubles in the caller's layout: pot_atm[class][atom] and the
// Fortran (3,3,natom,nregion) array per class
template <class AH>
__global__ __launch_bounds__(256) void k_hf_export(int natom, int nreg, int nw,
                                                   const AH *__restrict__ pe,
                                                   const AH *__restrict__ w,
                                                   double *__restrict__ pe_out,
                                                   double *__restrict__ w_out) {
  const size_t per_cls = (size_t)nreg * natom;
  const size_t total = (size_t)PGK_NHF * per_cls;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (size_t)gridDim.x * blockDim.x) {
    // k = (cls, region, atom); 9 outputs each
    double *o = w_out + 9 * k;
    const AH *src = w + (size_t)nw * k;
    if (nw == 9) {
      for (int q = 0; q < 9; q++) o[q] = acc_load(&src[q]);
    } else {
      for (int q = 0; q < 6; q++) o[q] = 0.0;
      for (int a = 0; a < 3; a++) o[6 + a] = acc_load(&src[a]);
    }
    if (k < (size_t)PGK_NHF * natom) pe_out[k] = acc_load(&pe[k]);
  }
}

template <int PREC>
int launch_hf(pgk_ctx *c, bool do_pairs) {
  using AH = typename HfAcc<PREC>::type;
  hipStream_t st = c->stream;
  HfDev h = c->hfdev;
  if (do_pairs && c->ncl > 0) {
    HfPairArgs p;
    p.ci_start = c->d_ci_start;
    p.cj = c->d_cj;
    p.cl_atom = c->d_cl_atom;
    p.mask = c->d_mask;
    p.witem = c->d_witem;
    p.ncl = c->ncl;
    p.nwi = c->nwi;
    p.ntype = c->ntype;
    double sq = c->rrcut * c->rrcut, sq2 = c->rcut * c->rcut;
    p.sqcut = sq2 > sq ? sq2 : sq;
    p.sqcut2 = sq2;
    p.rrcut = c->rrcut;
    p.rrsq = c->rrsq_cut;
    // the block count sizes jpart (hf_setup) and the collect pass reads it
    const int nblk = pgk_grid_blocks(pgk_pair_waves(c), 4);
    const bool ch = pgk_pair_chunked(c);
    if (!h.flux) return pgk_hf_w_launch(c);   // per-atom arrays: k_hf_w.hip
    // the flux form reads its atoms' data per cluster slot
    const int nslot = c->ncl * PGK_CL;
    int rc;
    if ((rc = grow_dev((void **)&c->d_hfw_xq, &c->hfw_xq_bytes, sizeof(double4) * nslot)) ||
        (rc = grow_dev((void **)&c->d_hfw_cen, &c->hfw_cen_bytes, sizeof(double4) * nslot)) ||
        (rc = grow_dev((void **)&c->d_hfw_mol, &c->hfw_mol_bytes, sizeof(int32_t) * nslot)) ||
        (rc = grow_dev((void **)&c->d_hfw_v, &c->hfw_v_bytes, sizeof(double4) * nslot)))
      return rc;
    k_hf_pack<<<pgk_grid_blocks(nslot, 256), 256, 0, st>>>(nslot, c->d_cl_atom, h, c->d_q,
                                                             c->d_hfw_xq, c->d_hfw_cen,
                                                             c->d_hfw_mol, c->d_hfw_v);
    HIP_CHECK(hipGetLastError());
    p.xq = c->d_hfw_xq;
    p.cen = c->d_hfw_cen;
    p.v4 = c->d_hfw_v;
    p.mol = c->d_hfw_mol;
    p.ct = c->d_ct;
    p.lj = c->d_ljp64;
    // the pair energies' transport share summed in the pass (see flux_add)
    const bool dt = h.mode == PGK_HF_BULK && h.direct_ok && h.vel == h.tvel;
    c->hfdev.direct_tr = dt ? 1 : 0;
    if (h.mode == PGK_HF_BULK) {
      auto kf = ch ? (dt ? k_hf_pairs_flux<AH, true, true, true>
                         : k_hf_pairs_flux<AH, true, true, false>)
                   : (dt ? k_hf_pairs_flux<AH, true, false, true>
                         : k_hf_pairs_flux<AH, true, false, false>);
      kf<<<nblk, 256, 0, st>>>(p, h);
    } else {
      const size_t lds = sizeof(double) * 18 * (size_t)h.nreg;   // 4 wave slices + bounds
      auto kf = ch ? k_hf_pairs_flux<AH, false, true, false>
                   : k_hf_pairs_flux<AH, false, false, false>;
      const int lr = pgk_lds_limit((const void *)kf, lds);
      if (lr) return lr;
      kf<<<nblk, 256, lds, st>>>(p, h);
    }
    HIP_CHECK(hipGetLastError());
  }
  return PGK_OK;
}

template <int PREC>
int export_hf(pgk_ctx *c) {
  using AH = typename HfAcc<PREC>::type;
  const HfDev &h = c->hfdev;
  size_t items = (size_t)PGK_NHF * h.nreg * h.natom;
  int nb = (int)std::min<size_t>((items + 255) / 256, 65536);
  k_hf_export<AH><<<nb, 256, 0, c->stream>>>(h.natom, h.nreg, h.nw,
                                             reinterpret_cast<const AH *>(h.pe),
                                             reinterpret_cast<const AH *>(h.w), c->d_hf_pe_out,
                                             c->d_hf_w_out);
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

}  // namespace

int pgk_hf_launch_pairs(pgk_ctx *c, bool do_pairs) {
  switch (c->precision) {
    case PGK_PREC_MIXED: return launch_hf<PGK_PREC_MIXED>(c, do_pairs);
    case PGK_PREC_FP64: return launch_hf<PGK_PREC_FP64>(c, do_pairs);
    default: return launch_hf<PGK_PREC_FP32>(c, do_pairs);
  }
}

int pgk_hf_export(pgk_ctx *c) {
  switch (c->precision) {
    case PGK_PREC_MIXED: return export_hf<PGK_PREC_MIXED>(c);
    case PGK_PREC_FP64: return export_hf<PGK_PREC_FP64>(c);
    default: return export_hf<PGK_PREC_FP32>(c);
  }
}
