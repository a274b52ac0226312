// k_bonded.hip - bonded terms, one thread per term, FP64 in every form
// (forces accumulate in the context's accumulator type).
//
//   calbondp    calbond.F90:109-219     E = K (r - r0)^2, skipped if |K| < 1e-16
//   calanglp    calangl.F90:180-358     E = K (th - th0)^2, skipped if |K| < 1e-32,
//                                       virial from absolute coordinates
//   caltorsp    caltors.F90:130-460     E = (V/div)(1 + cos(n phi - delta)),
//                                       force through the Chebyshev form of calijkl
//   caltorsrbp  caltorsrb.F90:130-459   E = sum_k (-1)^k C_k cos^k phi (Horner)
//   calnon14p   calnon14.F90:249-521    scaled Coulomb + LJ on torsion end atoms
#include "pgk_hf.h"

namespace {

__device__ __forceinline__ void mimg(double *r, const double3 &L, const double3 &Li) {
  r[0] = r[0] - L.x * round(r[0] * Li.x);
  r[1] = r[1] - L.y * round(r[1] * Li.y);
  r[2] = r[2] - L.z * round(r[2] * Li.z);
}
__device__ __forceinline__ double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void cross3(const double *x, const double *y, double *o) {
  o[0] = x[1] * y[2] - x[2] * y[1];
  o[1] = x[2] * y[0] - x[0] * y[2];
  o[2] = x[0] * y[1] - x[1] * y[0];
}

// f == nullptr: the call wants no forces (a heat-flux sample)
template <class A>
__device__ __forceinline__ void addf(A *f, int i, const double *v, double s) {
  if (!f) return;
  acc_add(&f[3 * i + 0], s * v[0]);
  acc_add(&f[3 * i + 1], s * v[1]);
  acc_add(&f[3 * i + 2], s * v[2]);
}

struct BArgs {
  const double *pos;
  const double *par;
  double3 L, Li;
  int n;
  double *red;   // energy and virial records; nullptr: none (a heat-flux sample)
};

// rec: E, V, T(9 in Fortran (a,n) order)
template <class A, class AH, int HFM>
__device__ __forceinline__ void bonds_body(int bid, const BArgs &a,
                                           const int32_t *__restrict__ bl, int ok, int orr,
                                           A *fat, const HfDev &h) {
  double ev[11];
#pragma unroll
  for (int k = 0; k < 11; k++) ev[k] = 0.0;
  extern __shared__ __align__(8) unsigned char bl_lds[];
  AH *sj = reinterpret_cast<AH *>(bl_lds);   // flux form: the block's sums
  AH jb[3] = {(AH)0, (AH)0, (AH)0};   // bulk flux form: this thread's sums
  if (HFM == 3) hf_flux_block_begin_vol<AH>(h, sj);
  int t = bid * blockDim.x + threadIdx.x;
  if (t < a.n) {
    int i = bl[3 * t], j = bl[3 * t + 1], ty = bl[3 * t + 2];
    HfAt at[2];
    constexpr bool fx = HFM >= 2, fb = HFM == 2;   // flux form (bulk): atoms loaded up front
    if constexpr (fx) {
      hf_at_load<fb>(h, i, at[0]);
      hf_at_load<fb>(h, j, at[1]);
    }
    double K = a.par[ok + ty], r0 = a.par[orr + ty];
    if (fabs(K) >= 1.0e-16) {
      double r[3];
      for (int d = 0; d < 3; d++) r[d] = a.pos[3 * i + d] - a.pos[3 * j + d];
      mimg(r, a.L, a.Li);
      double ra = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      ev[0] = K * ((ra - r0) * (ra - r0));
      double fa = -2.0 * K * (ra - r0);
      double f[3] = {fa * r[0] / ra, fa * r[1] / ra, fa * r[2] / ra};
      addf(fat, i, f, 1.0);
      addf(fat, j, f, -1.0);
      if (HFM) {   // calbondp_hf (calbond_hf.F90:6-164)
        const double fj[3] = {-f[0], -f[1], -f[2]};
        hf_pe_add<AH>(h, PGK_HF_BO, i, 0.5 * ev[0]);
        hf_pe_add<AH>(h, PGK_HF_BO, j, 0.5 * ev[0]);
        if constexpr (fx) calvirihf_flux_at<AH, fb>(h, at[0], at[1], f, fj, 0.5, jb, sj);
        else calvirihf_dev<AH>(h, i, j, f, fj, 0.5, PGK_HF_BO);
      }
      ev[1] = dot3(f, r);
      for (int nn = 0; nn < 3; nn++)
        for (int aa = 0; aa < 3; aa++) ev[2 + aa + 3 * nn] = f[aa] * r[nn];
    }
  }
  if (HFM == 2) hf_flux_block_end_bulk<AH>(h, jb, sj, PGK_HF_BO);
  if (HFM == 3) hf_flux_block_end_vol<AH>(h, sj, PGK_HF_BO);
  if (a.red) block_record<11, 4>(ev, a.red + (size_t)bid * PGK_REC);
}

template <class A, class AH, int HFM>
__device__ __forceinline__ void angles_body(int bid, const BArgs &a,
                                            const int32_t *__restrict__ al, int ok, int ot,
                                            A *fat, const HfDev &h) {
  double ev[11];
#pragma unroll
  for (int k = 0; k < 11; k++) ev[k] = 0.0;
  extern __shared__ __align__(8) unsigned char bl_lds[];
  AH *sj = reinterpret_cast<AH *>(bl_lds);   // flux form: the block's sums
  AH jb[3] = {(AH)0, (AH)0, (AH)0};   // bulk flux form: this thread's sums
  if (HFM == 3) hf_flux_block_begin_vol<AH>(h, sj);
  int t = bid * blockDim.x + threadIdx.x;
  if (t < a.n) {
    int i = al[4 * t], j = al[4 * t + 1], k = al[4 * t + 2], ty = al[4 * t + 3];
    HfAt at[3];
    constexpr bool fx = HFM >= 2, fb = HFM == 2;   // flux form (bulk): atoms loaded up front
    if constexpr (fx) {
      hf_at_load<fb>(h, i, at[0]);
      hf_at_load<fb>(h, j, at[1]);
      hf_at_load<fb>(h, k, at[2]);
    }
    double K = a.par[ok + ty], t0 = a.par[ot + ty];
    if (fabs(K) >= 1.0e-32) {
      double rij[3], rkj[3];
      for (int d = 0; d < 3; d++) {
        rij[d] = a.pos[3 * j + d] - a.pos[3 * i + d];
        rkj[d] = a.pos[3 * j + d] - a.pos[3 * k + d];
      }
      mimg(rij, a.L, a.Li);
      mimg(rkj, a.L, a.Li);
      double ra = sqrt(dot3(rij, rij)), rb = sqrt(dot3(rkj, rkj));
      double ct = dot3(rij, rkj) / (ra * rb);
      ct = fmin(1.0, fmax(-1.0, ct));
      double th = acos(ct), sn = sin(th);
      double dth = th - t0;
      ev[0] = K * (dth * dth);
      double fi[3] = {0, 0, 0}, fj[3] = {0, 0, 0}, fk[3] = {0, 0, 0};
      if (fabs(sn) > 1.0e-14) {
        double ft = -2.0 * K * dth / sn;
        for (int d = 0; d < 3; d++) {
          fi[d] = ft * ((rkj[d] / rb - rij[d] * ct / ra) / ra);
          fk[d] = ft * ((rij[d] / ra - rkj[d] * ct / rb) / rb);
          fj[d] = -fi[d] - fk[d];
        }
      }
      addf(fat, i, fi, 1.0);
      addf(fat, j, fj, 1.0);
      addf(fat, k, fk, 1.0);
      if (HFM) {   // calanglp_hf (calangl_hf.F90:6-255): 1/3 per atom, three scatters
        const double c3 = 1.0 / 3.0;
        hf_pe_add<AH>(h, PGK_HF_AN, i, c3 * ev[0]);
        hf_pe_add<AH>(h, PGK_HF_AN, j, c3 * ev[0]);
        hf_pe_add<AH>(h, PGK_HF_AN, k, c3 * ev[0]);
        if constexpr (fx) {
          calvirihf_flux_at<AH, fb>(h, at[0], at[1], fi, fj, c3, jb, sj);
          calvirihf_flux_at<AH, fb>(h, at[1], at[2], fj, fk, c3, jb, sj);
          calvirihf_flux_at<AH, fb>(h, at[2], at[0], fk, fi, c3, jb, sj);
        } else {
          calvirihf_dev<AH>(h, i, j, fi, fj, c3, PGK_HF_AN);
          calvirihf_dev<AH>(h, j, k, fj, fk, c3, PGK_HF_AN);
          calvirihf_dev<AH>(h, k, i, fk, fi, c3, PGK_HF_AN);
        }
      }
      const double *pi = a.pos + 3 * i, *pj = a.pos + 3 * j, *pk = a.pos + 3 * k;
      ev[1] = dot3(fi, pi) + dot3(fj, pj) + dot3(fk, pk);
      for (int nn = 0; nn < 3; nn++)
        for (int aa = 0; aa < 3; aa++)
          ev[2 + aa + 3 * nn] = fi[aa] * pi[nn] + fj[aa] * pj[nn] + fk[aa] * pk[nn];
    }
  }
  if (HFM == 2) hf_flux_block_end_bulk<AH>(h, jb, sj, PGK_HF_AN);
  if (HFM == 3) hf_flux_block_end_vol<AH>(h, sj, PGK_HF_AN);
  if (a.red) block_record<11, 4>(ev, a.red + (size_t)bid * PGK_REC);
}

// shared torsion geometry (calijkl / calijklrb)
struct TGeo {
  double rij[3], rkj[3], rkl[3], rik[3], rjl[3];
  double ni[3], nj[3], nia, nja, c, phi, s;
};

__device__ __forceinline__ void tgeo(const double *pos, int i, int j, int k, int l,
                                     const double3 &L, const double3 &Li, TGeo &g) {
  for (int d = 0; d < 3; d++) {
    g.rij[d] = pos[3 * j + d] - pos[3 * i + d];
    g.rkj[d] = pos[3 * j + d] - pos[3 * k + d];
    g.rkl[d] = pos[3 * l + d] - pos[3 * k + d];
    g.rik[d] = pos[3 * k + d] - pos[3 * i + d];
    g.rjl[d] = pos[3 * l + d] - pos[3 * j + d];
  }
  mimg(g.rij, L, Li);
  mimg(g.rkj, L, Li);
  mimg(g.rkl, L, Li);
  mimg(g.rik, L, Li);
  mimg(g.rjl, L, Li);
  cross3(g.rij, g.rkj, g.ni);
  cross3(g.rkj, g.rkl, g.nj);
  g.nia = sqrt(dot3(g.ni, g.ni));
  g.nja = sqrt(dot3(g.nj, g.nj));
  double c = dot3(g.ni, g.nj) / (g.nia * g.nja);
  c = fmin(1.0, fmax(-1.0, c));
  double t[3];
  cross3(g.ni, g.nj, t);
  double phi = acos(c);
  phi = -copysign(phi, dot3(g.rkj, t));
  g.c = c;
  g.phi = phi;
  g.s = sin(phi);
}

__device__ __forceinline__ void tforces(TGeo &g, double fa, double *fi, double *fj,
                                        double *fk, double *fl) {
  double ni[3], nj[3], fb[3], fc[3], t1[3], t2[3];
  for (int d = 0; d < 3; d++) {
    ni[d] = g.ni[d] / g.nia;
    nj[d] = g.nj[d] / g.nja;
  }
  for (int d = 0; d < 3; d++) {
    fb[d] = (nj[d] - g.c * ni[d]) / g.nia;
    fc[d] = (ni[d] - g.c * nj[d]) / g.nja;
  }
  cross3(fb, g.rkj, t1);
  for (int d = 0; d < 3; d++) fi[d] = fa * t1[d];
  cross3(fc, g.rkl, t1);
  cross3(fb, g.rik, t2);
  for (int d = 0; d < 3; d++) fj[d] = fa * (-t1[d] + t2[d]);
  cross3(fb, g.rij, t1);
  cross3(fc, g.rjl, t2);
  for (int d = 0; d < 3; d++) fk[d] = fa * (-t1[d] + t2[d]);
  cross3(fc, g.rkj, t1);
  for (int d = 0; d < 3; d++) fl[d] = fa * t1[d];
}

// RB = true: C_0..C_5 at par[oc + 6*ty]; false: periodic (div, barhig, phase, period)
template <class A, class AH, int HFM, bool RB>
__device__ __forceinline__ void tors_body(int bid, const BArgs &a,
                                          const int32_t *__restrict__ tl, int o0, int o1, int o2,
                                          int o3, A *fat, const HfDev &h) {
  double ev[11];
#pragma unroll
  for (int k = 0; k < 11; k++) ev[k] = 0.0;
  extern __shared__ __align__(8) unsigned char bl_lds[];
  AH *sj = reinterpret_cast<AH *>(bl_lds);   // flux form: the block's sums
  AH jb[3] = {(AH)0, (AH)0, (AH)0};   // bulk flux form: this thread's sums
  if (HFM == 3) hf_flux_block_begin_vol<AH>(h, sj);
  int t = bid * blockDim.x + threadIdx.x;
  if (t < a.n) {
    int i = tl[5 * t], j = tl[5 * t + 1], k = tl[5 * t + 2], l = tl[5 * t + 3], ty = tl[5 * t + 4];
    HfAt at[4];
    constexpr bool fx = HFM >= 2, fb = HFM == 2;   // flux form (bulk): atoms loaded up front
    if constexpr (fx) {
      hf_at_load<fb>(h, i, at[0]);
      hf_at_load<fb>(h, j, at[1]);
      hf_at_load<fb>(h, k, at[2]);
      hf_at_load<fb>(h, l, at[3]);
    }
    TGeo g;
    tgeo(a.pos, i, j, k, l, a.L, a.Li, g);
    double fa, dp;
    if (RB) {
      const double *C = a.par + o0 + 6 * ty;
      double b[6] = {C[0], -C[1], C[2], -C[3], C[4], -C[5]};
      dp = b[5];
      fa = 0.0;
      for (int q = 4; q >= 0; q--) {
        fa = fa * g.c + dp;
        dp = dp * g.c + b[q];
      }
      fa = -fa;
    } else {
      double div = a.par[o0 + ty], bh = a.par[o1 + ty], ph = a.par[o2 + ty];
      double per = fabs(a.par[o3 + ty]);
      dp = bh / div * (1.0 + cos(per * g.phi - ph));
      double c = g.c;
      int np = (int)rint(per);
      switch (np) {
        case 1: fa = 1.0; break;
        case 2: fa = 2.0 * c; break;
        case 3: fa = -4.0 * (g.s * g.s) + 3.0; break;
        case 4: fa = 4.0 * c * (2.0 * (c * c) - 1.0); break;
        case 5: fa = 4.0 * (c * c); fa = fa * (fa - 3.0) + 1.0; break;
        default: fa = 4.0 * (c * c); fa = 2.0 * c * (fa - 3.0) * (fa - 1.0); break;
      }
      fa = -fa * cos(ph) * bh * per / div;
    }
    double fi[3], fj[3], fk[3], fl[3];
    tforces(g, fa, fi, fj, fk, fl);
    addf(fat, i, fi, 1.0);
    addf(fat, j, fj, 1.0);
    addf(fat, k, fk, 1.0);
    addf(fat, l, fl, 1.0);
    if (HFM) {   // caltors(rb)p_hf: 1/4 per atom, six pair scatters
      const double c4 = 0.25;
      hf_pe_add<AH>(h, PGK_HF_TO, i, c4 * dp);
      hf_pe_add<AH>(h, PGK_HF_TO, j, c4 * dp);
      hf_pe_add<AH>(h, PGK_HF_TO, k, c4 * dp);
      hf_pe_add<AH>(h, PGK_HF_TO, l, c4 * dp);
      if constexpr (fx) {
        calvirihf_flux_at<AH, fb>(h, at[0], at[1], fi, fj, c4, jb, sj);
        calvirihf_flux_at<AH, fb>(h, at[0], at[2], fi, fk, c4, jb, sj);
        calvirihf_flux_at<AH, fb>(h, at[0], at[3], fi, fl, c4, jb, sj);
        calvirihf_flux_at<AH, fb>(h, at[1], at[2], fj, fk, c4, jb, sj);
        calvirihf_flux_at<AH, fb>(h, at[1], at[3], fj, fl, c4, jb, sj);
        calvirihf_flux_at<AH, fb>(h, at[2], at[3], fk, fl, c4, jb, sj);
      } else {
        calvirihf_dev<AH>(h, i, j, fi, fj, c4, PGK_HF_TO);
        calvirihf_dev<AH>(h, i, k, fi, fk, c4, PGK_HF_TO);
        calvirihf_dev<AH>(h, i, l, fi, fl, c4, PGK_HF_TO);
        calvirihf_dev<AH>(h, j, k, fj, fk, c4, PGK_HF_TO);
        calvirihf_dev<AH>(h, j, l, fj, fl, c4, PGK_HF_TO);
        calvirihf_dev<AH>(h, k, l, fk, fl, c4, PGK_HF_TO);
      }
    }
    const double *P[4] = {a.pos + 3 * i, a.pos + 3 * j, a.pos + 3 * k, a.pos + 3 * l};
    const double *F[4] = {fi, fj, fk, fl};
    ev[0] = dp;
    for (int q = 0; q < 4; q++) ev[1] += dot3(F[q], P[q]);
    for (int nn = 0; nn < 3; nn++)
      for (int aa = 0; aa < 3; aa++) {
        double s = 0.0;
        for (int q = 0; q < 4; q++) s += F[q][aa] * P[q][nn];
        ev[2 + aa + 3 * nn] = s;
      }
  }
  if (HFM == 2) hf_flux_block_end_bulk<AH>(h, jb, sj, PGK_HF_TO);
  if (HFM == 3) hf_flux_block_end_vol<AH>(h, sj, PGK_HF_TO);
  if (a.red) block_record<11, 4>(ev, a.red + (size_t)bid * PGK_REC);
}

// 1-4 pairs; rec: E_elc14, E_vdw14, V, T(9)
template <class A, class AH, int HFM>
__device__ __forceinline__ void non14_body(int bid, const BArgs &a,
                                           const int32_t *__restrict__ pl,
                                           const double *__restrict__ pf,
                                           const double *__restrict__ q,
                                           const int32_t *__restrict__ type,
                                           const double *__restrict__ eps,
                                           const double *__restrict__ rmin, int ntype, A *fat,
                                           const HfDev &h) {
  double ev[12];
#pragma unroll
  for (int k = 0; k < 12; k++) ev[k] = 0.0;
  extern __shared__ __align__(8) unsigned char bl_lds[];
  AH *sj = reinterpret_cast<AH *>(bl_lds);   // flux form: the block's sums
  AH jb[3] = {(AH)0, (AH)0, (AH)0};   // bulk flux form: this thread's sums
  if (HFM == 3) hf_flux_block_begin_vol<AH>(h, sj);
  int t = bid * blockDim.x + threadIdx.x;
  if (t < a.n) {
    int i = pl[2 * t], j = pl[2 * t + 1];
    HfAt at[2];
    constexpr bool fx = HFM >= 2, fb = HFM == 2;   // flux form (bulk): atoms loaded up front
    if constexpr (fx) {
      hf_at_load<fb>(h, i, at[0]);
      hf_at_load<fb>(h, j, at[1]);
    }
    double dv = pf[2 * t], de = pf[2 * t + 1];
    int tp = (type[i] - 1) + ntype * (type[j] - 1);
    double r[3];
    for (int d = 0; d < 3; d++) r[d] = a.pos[3 * i + d] - a.pos[3 * j + d];
    mimg(r, a.L, a.Li);
    double r2 = dot3(r, r);
    double ri2 = 1.0 / r2;
    double ri1 = sqrt(ri2);
    double pel = q[i] * q[j] * ri1;
    pel = pel * de;
    double fe[3] = {pel * ri2 * r[0], pel * ri2 * r[1], pel * ri2 * r[2]};
    double rad = rmin[tp];
    double t2 = ri2 * (rad * rad), t6 = t2 * t2 * t2, t12 = t6 * t6;
    double e = eps[tp] * dv;
    double pv = e * (t12 - 2.0 * t6);
    double ft = 12.0 * e * (t12 - t6) * ri2;
    double f[3] = {fe[0] + ft * r[0], fe[1] + ft * r[1], fe[2] + ft * r[2]};
    addf(fat, i, f, 1.0);
    addf(fat, j, f, -1.0);
    if (HFM) {   // calnon14p_hf (calnon14_hf.F90:8-386): Coulomb and LJ parts apart
      const double fej[3] = {-fe[0], -fe[1], -fe[2]};
      const double fv[3] = {ft * r[0], ft * r[1], ft * r[2]};
      const double fvj[3] = {-fv[0], -fv[1], -fv[2]};
      hf_pe_add<AH>(h, PGK_HF_14, i, 0.5 * pel);
      hf_pe_add<AH>(h, PGK_HF_14, j, 0.5 * pel);
      hf_pe_add<AH>(h, PGK_HF_14, i, 0.5 * pv);
      hf_pe_add<AH>(h, PGK_HF_14, j, 0.5 * pv);
      if constexpr (fx) {
        calvirihf_flux_at<AH, fb>(h, at[0], at[1], fe, fej, 0.5, jb, sj);
        calvirihf_flux_at<AH, fb>(h, at[0], at[1], fv, fvj, 0.5, jb, sj);
      } else {
        calvirihf_dev<AH>(h, i, j, fe, fej, 0.5, PGK_HF_14);
        calvirihf_dev<AH>(h, i, j, fv, fvj, 0.5, PGK_HF_14);
      }
    }
    ev[0] = pel;
    ev[1] = pv;
    ev[2] = dot3(f, r);
    for (int nn = 0; nn < 3; nn++)
      for (int aa = 0; aa < 3; aa++) ev[3 + aa + 3 * nn] = f[aa] * r[nn];
  }
  if (HFM == 2) hf_flux_block_end_bulk<AH>(h, jb, sj, PGK_HF_14);
  if (HFM == 3) hf_flux_block_end_vol<AH>(h, sj, PGK_HF_14);
  if (a.red) block_record<12, 4>(ev, a.red + (size_t)bid * PGK_REC);
}

// Every bonded term of one call in one launch: blocks [0, g1) bonds, then
// angles, periodic torsions, RB torsions and 1-4 pairs up to g5; each block
// runs its term's body with its index within the term.
template <class A>
struct BondedLaunch {
  BArgs b, an, t, rb, p14;
  const int32_t *bl, *al, *tl, *rl, *pl;
  int ob_k, ob_r, oa_k, oa_t, ot_div, ot_b, ot_p, ot_per, orb_c;
  const double *p14f, *q, *eps, *rmin;
  const int32_t *type;
  int ntype;
  int g1, g2, g3, g4, g5;
  A *fat;
};

template <class A, class AH, int HFM>
__global__ __launch_bounds__(256) void k_bonded(BondedLaunch<A> L, HfDev h) {
  const int b = blockIdx.x;
  if (b < L.g1) {
    bonds_body<A, AH, HFM>(b, L.b, L.bl, L.ob_k, L.ob_r, L.fat, h);
  } else if (b < L.g2) {
    angles_body<A, AH, HFM>(b - L.g1, L.an, L.al, L.oa_k, L.oa_t, L.fat, h);
  } else if (b < L.g3) {
    tors_body<A, AH, HFM, false>(b - L.g2, L.t, L.tl, L.ot_div, L.ot_b, L.ot_p, L.ot_per, L.fat,
                                 h);
  } else if (b < L.g4) {
    tors_body<A, AH, HFM, true>(b - L.g3, L.rb, L.rl, L.orb_c, 0, 0, 0, L.fat, h);
  } else {
    non14_body<A, AH, HFM>(b - L.g4, L.p14, L.pl, L.p14f, L.q, L.type, L.eps, L.rmin, L.ntype,
                           L.fat, h);
  }
}

}  // namespace

// Launch every bonded term on the MTS level mts_flag.  Records are appended
// per term; the segment table (kind per segment) is written to segs.  With
// hf the calforce_hf variants also scatter the per-atom heat-flux terms.
// HFM: 0 forces only, 1 the per-atom heat-flux arrays (calforce_hf), 2 and 3
// the flux form (the block's flux sums) in the bulk and the volume mode; each
// form is its own kernel, so a flux form carries none of the others' code
template <int PREC, int HFM>
static int launch_bonded(pgk_ctx *c, const pgk_mts *m, int mts_flag, int *seg_off, int *seg_n,
                         int *seg_kind, int *nseg) {
  using A = typename Prec<PREC>::acc;
  using AH = typename HfAcc<PREC>::type;
  const HfDev &h = c->hfdev;
  // a heat-flux sample of the MD loop wants the per-atom energies and the
  // flux only: no forces, no records
  A *fat = c->hf_pairs_only ? nullptr : reinterpret_cast<A *>(c->d_fat);
  BArgs a;
  a.pos = c->d_pos;
  a.par = c->d_bpar;
  a.L = make_double3(c->box[0], c->box[1], c->box[2]);
  a.Li = make_double3(1.0 / c->box[0], 1.0 / c->box[1], 1.0 / c->box[2]);
  hipStream_t st = c->stream;
  // the flux form's per-block sums (at most 1024 regions: 8 KiB)
  const size_t lds = (HFM >= 2) ? sizeof(AH) * (size_t)hf_flux_block_len(h) : 0;
  auto seg = [&](int nblk, int kind) {
    seg_off[*nseg] = c->red_used;
    seg_n[*nseg] = nblk;
    seg_kind[*nseg] = kind;
    (*nseg)++;
    a.red = c->d_red + (size_t)c->red_used * PGK_REC;
    c->red_used += nblk;
    if (c->hf_pairs_only) a.red = nullptr;   // a heat-flux sample reads no records
  };
  BondedLaunch<A> L;
  L.fat = fat;
  int gend = 0;
  int nb = c->nbond_act;
  if (m->bond == mts_flag && nb > 0) {
    a.n = nb;
    const int g = pgk_grid_blocks(nb, 256);
    seg(g, 1);
    L.b = a;
    gend += g;
  }
  L.g1 = gend;
  L.bl = c->d_bond;
  L.ob_k = c->off_bk;
  L.ob_r = c->off_br;
  int na = c->nangl_act;
  if (m->angl == mts_flag && na > 0) {
    a.n = na;
    const int g = pgk_grid_blocks(na, 256);
    seg(g, 2);
    L.an = a;
    gend += g;
  }
  L.g2 = gend;
  L.al = c->d_angl;
  L.oa_k = c->off_ak;
  L.oa_t = c->off_at;
  int nt = (int)(c->tors_type.size());
  if (m->tors == mts_flag && nt > 0) {
    a.n = nt;
    const int g = pgk_grid_blocks(nt, 256);
    seg(g, 3);
    L.t = a;
    gend += g;
  }
  L.g3 = gend;
  L.tl = c->d_tors;
  L.ot_div = c->off_tdiv;
  L.ot_b = c->off_tb;
  L.ot_p = c->off_tp;
  L.ot_per = c->off_tper;
  int nr = (int)(c->rb_type.size());
  if (m->torsrb == mts_flag && nr > 0) {
    a.n = nr;
    const int g = pgk_grid_blocks(nr, 256);
    seg(g, 4);
    L.rb = a;
    gend += g;
  }
  L.g4 = gend;
  L.rl = c->d_rb;
  L.orb_c = c->off_rbc;
  if (m->nonb14 == mts_flag && c->n14 > 0) {
    a.n = c->n14;
    const int g = pgk_grid_blocks(c->n14, 256);
    seg(g, 5);
    L.p14 = a;
    gend += g;
  }
  L.g5 = gend;
  L.pl = c->d_p14;
  L.p14f = c->d_p14f;
  L.q = c->d_q;
  L.type = c->d_type;
  L.eps = c->d_eps;
  L.rmin = c->d_rmin;
  L.ntype = c->ntype;
  if (gend > 0) k_bonded<A, AH, HFM><<<gend, 256, lds, st>>>(L, h);
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

int pgk_bonded_launch2(pgk_ctx *c, const pgk_mts *m, int mts_flag, int *seg_off, int *seg_n,
                       int *seg_kind, int *nseg, bool hf) {
  const int hfm = !hf ? 0 : !c->hfdev.flux ? 1 : c->hfdev.mode == PGK_HF_BULK ? 2 : 3;
#define PGK_BL(P)                                                                          \
  return hfm == 3   ? launch_bonded<P, 3>(c, m, mts_flag, seg_off, seg_n, seg_kind, nseg)   \
         : hfm == 2 ? launch_bonded<P, 2>(c, m, mts_flag, seg_off, seg_n, seg_kind, nseg)   \
         : hfm == 1 ? launch_bonded<P, 1>(c, m, mts_flag, seg_off, seg_n, seg_kind, nseg)   \
                    : launch_bonded<P, 0>(c, m, mts_flag, seg_off, seg_n, seg_kind, nseg)
  switch (c->precision) {
    case PGK_PREC_FP64: PGK_BL(PGK_PREC_FP64);
    case PGK_PREC_MIXED: PGK_BL(PGK_PREC_MIXED);
    default: PGK_BL(PGK_PREC_FP32);
  }
#undef PGK_BL
}
