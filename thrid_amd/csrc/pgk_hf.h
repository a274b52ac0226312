// pgk_hf.h - per-atom heat-flux decomposition on the device (calforce_hf).
//
// Restates calvirihf (calvirihf.F90:8-405): for an interaction between atoms
// i and j with forces fi (on i) and fj (on j) and weight nbc, scatter
//   volume mode (default):  W(a,3,i,r) += fi_a ff,  W(a,3,j,r) += -fj_a ff,
//     ff = signed length of the z-segment between the two centres that lies
//     inside region r (segments with |dz| >= Lz/2 are skipped);
//   surface mode (MOP):     the same columns with ff = +-nbc for every region
//     plane z1(r) the segment crosses;
//   bulk mode:              W(a,n,i,r) += r_n fi_a nbc, W(a,n,j,r) += r_n (-fj_a) nbc
//     with the minimum-image r = c_i - c_j (all regions alike).
// Centres are atom positions (z wrapped once into [0, Lz) outside bulk mode) or
// molecule centres of mass for atoms with hftyp == MOLE; two MOLE atoms of the
// same molecule contribute nothing.
//
// Device layout: pe[class][atom]; w[class][region][atom][nw] with nw = 9
// (bulk: index a + 3n) or nw = 3 (volume/surface: column n = 3 only, index a).
#pragma once

#include "pgk_device.h"

// accumulator type of the heat-flux sums: int64 fixed point in the mixed form
// (order-independent), FP64 otherwise
template <int PREC> struct HfAcc { using type = double; };
template <> struct HfAcc<PGK_PREC_MIXED> { using type = unsigned long long; };

template <class AH>
__device__ __forceinline__ void hf_pe_add(const HfDev &h, int cls, int i, double v) {
  acc_add(reinterpret_cast<AH *>(h.pe) + (size_t)cls * h.natom + i, v);
}

__device__ __forceinline__ int hf_centre(const HfDev &h, int i, bool bulk, double *c) {
  if (h.hftyp[i] == PGK_HFTYP_MOLE) {
    const double *m = h.molecom + 3 * (size_t)(h.molid[i] - 1);
    c[0] = m[0];
    c[1] = m[1];
    c[2] = m[2];
    return 0;
  }
  c[0] = h.pos[3 * i];
  c[1] = h.pos[3 * i + 1];
  c[2] = h.pos[3 * i + 2];
  if (!bulk) {
    if (c[2] < 0.0) c[2] = c[2] + h.L[2];
    else if (c[2] >= h.L[2]) c[2] = c[2] - h.L[2];
  }
  return 1;
}

// The weight of region ir for a segment between centres with z coordinates
// ciz, cjz (r2 = ciz - cjz, rz = |r2|) in the volume form: the signed length
// of the part of the segment inside [z1, z2], times nbc (calvirihf.F90:201-357).
// Returns 0 when the segment misses the region.
__device__ __forceinline__ double hf_volume_weight_z(double z1, double z2, double ciz,
                                                     double cjz, double r2, double rz,
                                                     double nbc) {
  const double ri1 = ciz - z1, rj1 = cjz - z1, ri2 = ciz - z2, rj2 = cjz - z2;
  int np = 0;
  bool p1 = false;
  if (ri1 * rj1 < 0.0) {
    p1 = true;
    np++;
  }
  if (ri2 * rj2 < 0.0) np++;
  double ff;
  if (np == 0) {
    if (ri1 * ri2 > 0.0) return 0.0;
    ff = rz;
  } else if (np == 1) {
    if (p1) ff = (ri1 > 0.0) ? ri1 : rj1;
    else ff = (ri2 < 0.0) ? -ri2 : -rj2;
  } else {
    ff = z2 - z1;
  }
  return (r2 >= 0.0) ? ff * nbc : -ff * nbc;
}

__device__ __forceinline__ double hf_volume_weight(const HfDev &h, int ir, double ciz,
                                                   double cjz, double r2, double rz,
                                                   double nbc) {
  return hf_volume_weight_z(h.z1[ir], h.z2[ir], ciz, cjz, r2, rz, nbc);
}

// the regions a z segment [lo, hi] can touch: with equal bins the bins of its
// ends (widened by one for the rounding of the product by 1/width and the
// bins' deviation from equal; the exact weight decides), else all
__device__ __forceinline__ void hf_region_window(const HfDev &h, double lo, double hi, int &r0,
                                                 int &r1) {
  if (h.ubin) {
    r0 = max(0, (int)floor((lo - h.ubz0) * h.ubwi) - 1);
    r1 = min(h.nreg - 1, (int)floor((hi - h.ubz0) * h.ubwi) + 1);
  } else {
    r0 = 0;
    r1 = h.nreg - 1;
  }
}

// Flux form of calvirihf: the interaction's share of sum_a W(a,n,.,r) v(a),
// i.e. the region weight times s = fi . v_i - fj . v_j, in the heat-flux
// accumulator type.  Bulk mode (the same 3 components for every region):
// each thread sums its terms in registers (jb), the block adds its waves'
// sums once (hf_flux_block_end_bulk).  An add of every lane to one address
// would be serialised lane by lane.  Volume mode: the block's z components
// per region in LDS (sj), added to with LDS atomics (the lanes' regions
// differ) and flushed once per block (hf_flux_block_end_vol).  The pair pass
// keeps its own per-block sums.

// one term in the accumulator type: the fixed-point value of the mixed form
// is the one acc_add would add, so the sums are bit for bit the same
template <class AH> __device__ __forceinline__ AH acc_val(double v);
template <> __device__ __forceinline__ double acc_val<double>(double v) { return v; }
template <> __device__ __forceinline__ unsigned long long acc_val<unsigned long long>(double v) {
  return (unsigned long long)__double2ll_rn(v * PGK_FX_SCALE);
}

template <class AH, bool BULK>
__device__ __forceinline__ void calvirihf_flux_s(const HfDev &h, double s, const double *ci,
                                                 const double *cj, double nbc, AH *jb,
                                                 AH *sj) {
  if (BULK) {
    double r[3];
    for (int d = 0; d < 3; d++) {
      r[d] = ci[d] - cj[d];
      r[d] = r[d] - h.L[d] * round(r[d] * h.Li[d]);
    }
    for (int n = 0; n < 3; n++) jb[n] += acc_val<AH>(r[n] * nbc * s);
    return;
  }
  const double r2 = ci[2] - cj[2], rz = fabs(r2);
  if (rz * h.Li[2] >= 0.5) return;
  int r0, r1;
  hf_region_window(h, fmin(ci[2], cj[2]), fmax(ci[2], cj[2]), r0, r1);
  for (int ir = r0; ir <= r1; ir++) {
    const double ff = hf_volume_weight(h, ir, ci[2], cj[2], r2, rz, nbc);
    if (ff != 0.0) acc_add(&sj[ir], ff * s);
  }
}

// What the flux form needs of one atom of a bonded term: its centre,
// velocity, molecule and centre kind.  A bonded kernel loads all of a term's
// atoms at once, before the term's geometry, so that the loads' latency
// overlaps the arithmetic instead of following it pair by pair.
struct HfAt {
  double c[3], v[3];
  int mol, atom;   // atom: 1 for an ATOM centre, 0 for a MOLE centre
};

template <bool BULK>
__device__ __forceinline__ void hf_at_load(const HfDev &h, int i, HfAt &a) {
  a.mol = h.molid[i];
  a.atom = hf_centre(h, i, BULK, a.c);
  a.v[0] = h.vel[3 * i];
  a.v[1] = h.vel[3 * i + 1];
  a.v[2] = h.vel[3 * i + 2];
}

// calvirihf's flux form on loaded atoms (BULK: the bulk mode, else volume)
template <class AH, bool BULK>
__device__ __forceinline__ void calvirihf_flux_at(const HfDev &h, const HfAt &A, const HfAt &B,
                                                  const double *fi, const double *fj,
                                                  double nbc, AH *jb, AH *sj) {
  if (A.atom + B.atom == 0 && A.mol == B.mol) return;
  const double s = (fi[0] * A.v[0] + fi[1] * A.v[1] + fi[2] * A.v[2]) -
                   (fj[0] * B.v[0] + fj[1] * B.v[1] + fj[2] * B.v[2]);
  calvirihf_flux_s<AH, BULK>(h, s, A.c, B.c, nbc, jb, sj);
}

// entries of a block's flux sums in LDS (the dynamic LDS a bonded kernel
// needs in the flux form is hf_flux_block_len * sizeof(accumulator)): bulk,
// 3 per wave of a 256-thread block; volume, one per region
__host__ __device__ __forceinline__ int hf_flux_block_len(const HfDev &h) {
  return h.mode == PGK_HF_BULK ? 12 : h.nreg;
}

__device__ __forceinline__ void acc_add_raw(double *p, double v) { atomicAdd(p, v); }
__device__ __forceinline__ void acc_add_raw(unsigned long long *p, unsigned long long v) {
  atomicAdd(p, v);
}

// volume mode: every thread of the block calls both (they synchronise it)
template <class AH>
__device__ __forceinline__ void hf_flux_block_begin_vol(const HfDev &h, AH *sj) {
  for (int k = threadIdx.x; k < h.nreg; k += blockDim.x) sj[k] = (AH)0;
  __syncthreads();
}

template <class AH>
__device__ __forceinline__ void hf_flux_block_end_vol(const HfDev &h, const AH *sj, int cls) {
  __syncthreads();
  AH *J = reinterpret_cast<AH *>(h.jcol) + (size_t)cls * h.nreg * 3;
  for (int k = threadIdx.x; k < h.nreg; k += blockDim.x) {
    const AH v = sj[k];
    if (v != (AH)0) acc_add_raw(&J[3 * k + 2], v);
  }
}

// bulk mode: the threads' register sums jb over the wave (butterfly), the
// 4 waves in order, then one add per component and region; every thread of
// the 256-thread block calls it
template <class AH>
__device__ __forceinline__ void hf_flux_block_end_bulk(const HfDev &h, AH (&jb)[3], AH *sj,
                                                       int cls) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int n = 0; n < 3; n++)
    for (int o = 32; o >= 1; o >>= 1) jb[n] += __shfl_xor(jb[n], o, 64);
  if (lane == 0)
    for (int n = 0; n < 3; n++) sj[3 * wave + n] = jb[n];
  __syncthreads();
  AH *J = reinterpret_cast<AH *>(h.jcol) + (size_t)cls * h.nreg * 3;
  for (int k = threadIdx.x; k < 3 * h.nreg; k += blockDim.x) {
    const int n = k % 3;
    const AH v = ((sj[n] + sj[3 + n]) + sj[6 + n]) + sj[9 + n];
    if (v != (AH)0) acc_add_raw(&J[k], v);
  }
}

// the per-atom arrays (calforce_hf); the flux form is calvirihf_flux_at
template <class AH>
__device__ void calvirihf_dev(const HfDev &h, int i, int j, const double *fi, const double *fj,
                              double nbc, int cls) {
  const bool bulk = (h.mode == PGK_HF_BULK);
  double ci[3], cj[3];
  const int atomflag = hf_centre(h, i, bulk, ci) + hf_centre(h, j, bulk, cj);
  if (atomflag == 0 && h.molid[i] == h.molid[j]) return;
  double r[3] = {0.0, 0.0, 0.0}, rz = 0.0;
  if (bulk) {
    r[0] = ci[0] - cj[0];
    r[1] = ci[1] - cj[1];
  }
  r[2] = ci[2] - cj[2];
  if (!bulk) {
    rz = fabs(r[2]);
    if (rz * h.Li[2] >= 0.5) return;
  }
  AH *W = reinterpret_cast<AH *>(h.w) + (size_t)cls * h.nreg * h.natom * h.nw;
  const size_t N = (size_t)h.natom;
  if (h.mode == PGK_HF_SURFACE) {
    for (int ir = 0; ir < h.nreg; ir++) {
      const double ri1 = ci[2] - h.z1[ir], rj1 = cj[2] - h.z1[ir];
      if (ri1 * rj1 < 0.0) {
        double ff = (ri1 > 0.0) ? 1.0 : -1.0;
        ff = ff * nbc;
        AH *wi = W + (ir * N + i) * 3, *wj = W + (ir * N + j) * 3;
        for (int a = 0; a < 3; a++) {
          acc_add(&wi[a], fi[a] * ff);
          acc_add(&wj[a], (-fj[a]) * ff);
        }
      }
    }
    return;
  }
  if (bulk) {
    for (int d = 0; d < 3; d++) r[d] = r[d] - h.L[d] * round(r[d] * h.Li[d]);
    for (int ir = 0; ir < h.nreg; ir++) {
      AH *wi = W + (ir * N + i) * 9, *wj = W + (ir * N + j) * 9;
      for (int n = 0; n < 3; n++)
        for (int a = 0; a < 3; a++) {
          const double rfi = r[n] * fi[a], rfj = r[n] * (-fj[a]);
          acc_add(&wi[a + 3 * n], rfi * nbc);
          acc_add(&wj[a + 3 * n], rfj * nbc);
        }
    }
    return;
  }
  int r0, r1;
  hf_region_window(h, fmin(ci[2], cj[2]), fmax(ci[2], cj[2]), r0, r1);
  for (int ir = r0; ir <= r1; ir++) {
    const double ff = hf_volume_weight(h, ir, ci[2], cj[2], r[2], rz, nbc);
    if (ff == 0.0) continue;
    AH *wi = W + (ir * N + i) * 3, *wj = W + (ir * N + j) * 3;
    for (int a = 0; a < 3; a++) {
      acc_add(&wi[a], fi[a] * ff);
      acc_add(&wj[a], (-fj[a]) * ff);
    }
  }
}
