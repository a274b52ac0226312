// k_ewk.hip - the standard Ewald reciprocal sum with per-atom energy and
// virial: prepewk's wave vectors (prepewk.F90:80-150) and ewaldkp_hf
// (ewaldk_hf.F90:8-304), the calforce_hf path of the _HF_BULK_EWK build, for
// small systems (O(natom x nwave)).
//
// Per wave vector K (with A(K) = exp(-(pi/alpha)^2 |K|^2) / |K|^2):
//   S_s = sum_i q_i sin(2 pi K.r_i), S_c likewise, in atom order;
//   E_K = A (S_s^2 + S_c^2) / (2 pi V);
//   e_i += A q_i (sin_i S_s + cos_i S_c) / (2 pi V);
//   f_i += A (S_c sin_i - S_s cos_i) K, times 2 q_i / V at the end;
//   for K != 0 the tensor k_fact(a,n) = delta_an - (2/|K|^2 + 2 pi^2/alpha^2)
//   K_a K_n times E_K (box) and times the atom's e_i share (per atom), minus
//   A netchrgsq / (2 pi V) in both with the net-charge correction.
// Kernels: one thread per K sums S_s, S_c over the atoms in order
// (k_ewk_struct); one thread per atom walks the K in order for its energy,
// force and virial (k_ewk_atoms); one thread adds the box totals over K in
// order (k_ewk_total).  Every sum is in the reference's order; only the
// device sin / cos and contraction differ from it.
#include <cmath>

#include "pgk_internal.h"

namespace {

__global__ __launch_bounds__(256) void k_ewk_struct(int nw, int n, const double *__restrict__ kw,
                                                    const double *__restrict__ pos,
                                                    const double *__restrict__ q,
                                                    double *__restrict__ sc) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nw) return;
  const double pi2 = 2.0 * 3.141592653589793;
  const double kx = kw[3 * k], ky = kw[3 * k + 1], kz = kw[3 * k + 2];
  double bs = 0.0, bc = 0.0;
  for (int i = 0; i < n; i++) {
    double th = kx * pos[3 * i] + ky * pos[3 * i + 1] + kz * pos[3 * i + 2];
    th = pi2 * th;
    double s, c;
    sincos(th, &s, &c);
    bs = bs + q[i] * s;
    bc = bc + q[i] * c;
  }
  sc[2 * k] = bs;
  sc[2 * k + 1] = bc;
}

struct EwkArgs {
  int nw, n, netq;
  double pot_fact, inv_vol, fac_vir, alpha_sqrtpi, netchrgsq;
};

// k_fact(a, n) at a + 3n (ewaldk_hf.F90:200-212)
__device__ __forceinline__ void k_fact(const double *K, double fac_vir, double (&kf)[9]) {
#pragma clang fp contract(off)
  const double k2 = K[0] * K[0] + K[1] * K[1] + K[2] * K[2];
#pragma unroll
  for (int nn = 0; nn < 3; nn++)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      kf[a + 3 * nn] = -1.0 * (2.0 / k2 + fac_vir) * K[a] * K[nn];
      if (a == nn) kf[a + 3 * nn] = kf[a + 3 * nn] + 1.0;
    }
}

__global__ __launch_bounds__(64) void k_ewk_atoms(EwkArgs a, const double *__restrict__ kw,
                                                  const double *__restrict__ aw,
                                                  const double *__restrict__ sc,
                                                  const double *__restrict__ pos,
                                                  const double *__restrict__ q,
                                                  double *__restrict__ f_out,
                                                  double *__restrict__ e_out,
                                                  double *__restrict__ w_out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const double pi2 = 2.0 * 3.141592653589793;
  const double r0 = pos[3 * i], r1 = pos[3 * i + 1], r2 = pos[3 * i + 2], qi = q[i];
  double e = 0.0, f[3] = {0.0, 0.0, 0.0}, w[9] = {};
  for (int k = 0; k < a.nw; k++) {
    const double K[3] = {kw[3 * k], kw[3 * k + 1], kw[3 * k + 2]};
    double th = K[0] * r0 + K[1] * r1 + K[2] * r2;
    th = pi2 * th;
    double s, c;
    sincos(th, &s, &c);
    const double bs = sc[2 * k], bc = sc[2 * k + 1];
    const double pfa = a.pot_fact * aw[k];
    const double share = pfa * qi * (s * bs + c * bc);
    e = e + share;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double csk = aw[k] * bs * K[d], cck = aw[k] * bc * K[d];
      f[d] = f[d] + cck * s - csk * c;
    }
    if (k == 0) continue;
    double kf[9];
    k_fact(K, a.fac_vir, kf);
    const double fv3 = a.netq ? pfa * a.netchrgsq : 0.0;
#pragma unroll
    for (int t = 0; t < 9; t++) w[t] = w[t] + kf[t] * (share - fv3);
  }
  for (int d = 0; d < 3; d++) f_out[3 * i + d] = 2.0 * qi * a.inv_vol * f[d];
  e_out[i] = e - a.alpha_sqrtpi * qi * qi;
  for (int t = 0; t < 9; t++) w_out[(size_t)9 * i + t] = w[t];
}

// the box totals: pot_ewk, the net-charge term and the tensor, over K in order
__global__ void k_ewk_total(EwkArgs a, const double *__restrict__ kw,
                            const double *__restrict__ aw, const double *__restrict__ sc,
                            double *__restrict__ tot) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double ewk = 0.0, cc = 0.0, t9[9] = {};
  for (int k = 0; k < a.nw; k++) {
    const double bs = sc[2 * k], bc = sc[2 * k + 1];
    const double pfa = a.pot_fact * aw[k];
    const double poti = pfa * (bs * bs + bc * bc);
    ewk = ewk + poti;
    if (k == 0) continue;
    const double K[3] = {kw[3 * k], kw[3 * k + 1], kw[3 * k + 2]};
    double kf[9];
    k_fact(K, a.fac_vir, kf);
    double fv3 = 0.0;
    if (a.netq) {
      fv3 = pfa * a.netchrgsq;
      const double k2 = K[0] * K[0] + K[1] * K[1] + K[2] * K[2];
      cc = cc + fv3 * (-1.0 + a.fac_vir * k2);
    }
    for (int t = 0; t < 9; t++) t9[t] = t9[t] + kf[t] * (poti - fv3);
  }
  tot[0] = ewk;
  tot[1] = cc;
  for (int t = 0; t < 9; t++) tot[2 + t] = t9[t];
}

}  // namespace

extern "C" int pgk_ewaldk_hf(pgk_ctx *c, int32_t kmax, double yratio, double zratio,
                             int32_t ifnetq, double *force, double *pot_ewk,
                             double *pot_viri_coul, double *virt, double *pot_atm,
                             double *viri_atm) {
  if (!c || !force || !pot_ewk || !pot_viri_coul || !virt || !pot_atm || !viri_atm ||
      kmax < 1 || !(yratio > 0) || !(zratio > 0)) {
    pgk_set_error("pgk_ewaldk_hf: null argument or kmax < 1");
    return PGK_EINVAL;
  }
  if (!c->have_pos) {
    pgk_set_error("pgk_ewaldk_hf: positions not set");
    return PGK_ESTATE;
  }
  // wave vectors (prepewk.F90:80-150); entry 0 is K = 0 with A = 0
  const int ry = (int)std::round(yratio), rz = (int)std::round(zratio);
  if (ry < 1 || rz < 1) {
    pgk_set_error("pgk_ewaldk_hf: y and z ratios must round to at least 1");
    return PGK_EINVAL;
  }
  const double pi = std::acos(-1.0);
  std::vector<double> kw(3, 0.0), aw(1, 0.0);
  const int ykmax = kmax * ry, zkmax = kmax * rz, kmax2 = kmax * kmax;
  const double coeff = (pi / c->alpha) * (pi / c->alpha);
  for (int kx = -kmax; kx <= kmax; kx++)
    for (int ky = -ykmax; ky <= ykmax; ky++)
      for (int kz = -zkmax; kz <= zkmax; kz++) {
        if (kx == 0 && ky == 0 && kz == 0) continue;
        const double k2 = (double)(kx * kx) + (double)(ky * ky) / (double)(ry * ry) +
                          (double)(kz * kz) / (double)(rz * rz);
        if (!(k2 <= (double)kmax2)) continue;
        const double nv[3] = {(double)kx, (double)ky, (double)kz};
        const double kk = (nv[0] / c->box[0]) * (nv[0] / c->box[0]) +
                          (nv[1] / c->box[1]) * (nv[1] / c->box[1]) +
                          (nv[2] / c->box[2]) * (nv[2] / c->box[2]);
        aw.push_back(std::exp(-coeff * kk) / kk);
        for (int d = 0; d < 3; d++) kw.push_back(nv[d] / c->box[d]);
      }
  const int nw = (int)aw.size(), n = c->natom;
  HIP_CHECK(hipSetDevice(c->device));
  double *d_kw = nullptr, *d_aw = nullptr, *d_sc = nullptr, *d_out = nullptr;
  const size_t out_n = (size_t)3 * n + n + (size_t)9 * n + 11;
  auto release = [&]() {
    for (double *p : {d_kw, d_aw, d_sc, d_out})
      if (p) (void)hipFree(p);
  };
  if (hipMalloc((void **)&d_kw, sizeof(double) * kw.size()) != hipSuccess ||
      hipMalloc((void **)&d_aw, sizeof(double) * nw) != hipSuccess ||
      hipMalloc((void **)&d_sc, sizeof(double) * 2 * nw) != hipSuccess ||
      hipMalloc((void **)&d_out, sizeof(double) * out_n) != hipSuccess) {
    release();
    pgk_set_error("pgk_ewaldk_hf: device allocation failed");
    return PGK_ENOMEM;
  }
  hipStream_t st = c->stream;
  EwkArgs a;
  a.nw = nw;
  a.n = n;
  a.netq = ifnetq != 0;
  const double vol = c->box[0] * c->box[1] * c->box[2];
  a.pot_fact = 0.5 / pi / vol;
  a.inv_vol = 1.0 / vol;
  a.fac_vir = 2.0 * pi * pi * (1.0 / (c->alpha * c->alpha));
  a.alpha_sqrtpi = c->alpha / std::sqrt(pi);
  a.netchrgsq = c->netchrgsq;
  double *f_d = d_out, *e_d = d_out + 3 * n, *w_d = d_out + 4 * n, *t_d = d_out + 13 * (size_t)n;
  hipError_t e = hipMemcpyAsync(d_kw, kw.data(), sizeof(double) * kw.size(),
                                hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_aw, aw.data(), sizeof(double) * nw, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    k_ewk_struct<<<(nw + 255) / 256, 256, 0, st>>>(nw, n, d_kw, c->d_pos, c->d_q, d_sc);
    k_ewk_atoms<<<(n + 63) / 64, 64, 0, st>>>(a, d_kw, d_aw, d_sc, c->d_pos, c->d_q, f_d, e_d,
                                               w_d);
    k_ewk_total<<<1, 64, 0, st>>>(a, d_kw, d_aw, d_sc, t_d);
    e = hipGetLastError();
  }
  std::vector<double> h(out_n);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), d_out, sizeof(double) * out_n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  release();
  if (e != hipSuccess) {
    pgk_set_error(std::string("pgk_ewaldk_hf: ") + hipGetErrorString(e));
    return PGK_EDEVICE;
  }
  // the reference's accumulation into the caller's arrays
  for (int i = 0; i < 3 * n; i++) force[i] = force[i] + h[i];
  for (int i = 0; i < n; i++) pot_atm[i] = pot_atm[i] + h[3 * n + i];
  for (size_t i = 0; i < (size_t)9 * n; i++) viri_atm[i] = viri_atm[i] + h[4 * (size_t)n + i];
  const double *t = h.data() + 13 * (size_t)n;
  *pot_ewk = t[0];
  *pot_viri_coul = *pot_viri_coul + t[0] + t[1];
  for (int k = 0; k < 9; k++) virt[k] = virt[k] + t[2 + k];
  return PGK_OK;
}
