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
__device__ __forceinline__ double hf_volume_weight(const HfDev &h, int ir, double ciz,
                                                   double cjz, double r2, double rz,
                                                   double nbc) {
  const double z1 = h.z1[ir], z2 = h.z2[ir];
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

// the regions a z segment [lo, hi] can touch: with equal bins the bins of its
// ends (widened by one for rounding; the exact weight decides), else all
__device__ __forceinline__ void hf_region_window(const HfDev &h, double lo, double hi, int &r0,
                                                 int &r1) {
  if (h.ubin) {
    r0 = max(0, (int)floor((lo - h.ubz0) / h.ubw) - 1);
    r1 = min(h.nreg - 1, (int)floor((hi - h.ubz0) / h.ubw) + 1);
  } else {
    r0 = 0;
    r1 = h.nreg - 1;
  }
}

// flux form of calvirihf: the interaction's share of sum_a W(a,n,.,r) v(a),
// i.e. the region weight times s = fi . v_i - fj . v_j, added to
// jcol[cls][r][n] (global FP64 atomics; the bonded terms use this, the pair
// pass keeps its own per-block sums)
__device__ void calvirihf_flux(const HfDev &h, int i, int j, const double *fi, const double *fj,
                               double nbc, int cls, const double *ci, const double *cj) {
  const bool bulk = (h.mode == PGK_HF_BULK);
  const double *vi = h.vel + 3 * i, *vj = h.vel + 3 * j;
  const double s = (fi[0] * vi[0] + fi[1] * vi[1] + fi[2] * vi[2]) -
                   (fj[0] * vj[0] + fj[1] * vj[1] + fj[2] * vj[2]);
  double *J = h.jcol + (size_t)cls * h.nreg * 3;
  if (bulk) {
    double r[3];
    for (int d = 0; d < 3; d++) {
      r[d] = ci[d] - cj[d];
      r[d] = r[d] - h.L[d] * round(r[d] * h.Li[d]);
    }
    for (int ir = 0; ir < h.nreg; ir++)
      for (int n = 0; n < 3; n++) atomicAdd(&J[3 * ir + n], r[n] * nbc * s);
    return;
  }
  const double r2 = ci[2] - cj[2], rz = fabs(r2);
  if (rz * h.Li[2] >= 0.5) return;
  int r0, r1;
  hf_region_window(h, fmin(ci[2], cj[2]), fmax(ci[2], cj[2]), r0, r1);
  for (int ir = r0; ir <= r1; ir++) {
    const double ff = hf_volume_weight(h, ir, ci[2], cj[2], r2, rz, nbc);
    if (ff != 0.0) atomicAdd(&J[3 * ir + 2], ff * s);
  }
}

template <class AH>
__device__ void calvirihf_dev(const HfDev &h, int i, int j, const double *fi, const double *fj,
                              double nbc, int cls) {
  const bool bulk = (h.mode == PGK_HF_BULK);
  double ci[3], cj[3];
  const int atomflag = hf_centre(h, i, bulk, ci) + hf_centre(h, j, bulk, cj);
  if (atomflag == 0 && h.molid[i] == h.molid[j]) return;
  if (h.flux) {
    calvirihf_flux(h, i, j, fi, fj, nbc, cls, ci, cj);
    return;
  }
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
