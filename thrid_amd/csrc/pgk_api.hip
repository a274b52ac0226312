// pgk_api.hip - C ABI of libpgkforce.so (declared in include/pgk_force.h).
//
// Host-side set-up restates the reference's one-time preparation:
//   spline_interp_init / cal_spl_coeff  spline_interp.F90:14-190
//   fft_pme_init (moduli, self energy)  fft_pme_init.F90:14-182, 263-289
//   erf_corr_cutoff (rkcut)             erf_corr_cutoff.F90:28-62
//   netchrgsq                           prepmd.F90:244-249
// and pgk_calforce orchestrates the device kernels in calforce's order with
// its output semantics (calforce.F90:307-1068).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "pgk_device.h"

static thread_local std::string g_err;
void pgk_set_error(const std::string &msg) { g_err = msg; }

int pgk_lds_limit(const void *f, size_t lds) {
  static std::mutex mu;
  static std::map<std::pair<int, const void *>, size_t> set;
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  size_t &cur = set[std::make_pair(dev, f)];
  if (lds <= cur) return PGK_OK;
  HIP_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  cur = lds;
  return PGK_OK;
}

void pgk_prof_begin(pgk_ctx *c, int k) {
  if (!c->prof || !((c->prof_mask >> k) & 1u)) return;
  const int i = c->pn++;
  while ((int)c->pev.size() < 2 * (i + 1)) {
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    c->pev.push_back(e);
  }
  if ((int)c->pstage.size() < i + 1) c->pstage.resize(i + 1);
  c->pstage[i] = k;
  c->popen[k] = i;
  (void)hipEventRecord(c->pev[2 * i], c->stream);
}
void pgk_prof_end(pgk_ctx *c, int k) {
  if (!c->prof || !((c->prof_mask >> k) & 1u)) return;
  (void)hipEventRecord(c->pev[2 * c->popen[k] + 1], c->stream);
}
// call after the stream has been synchronised
void pgk_prof_collect(pgk_ctx *c) {
  for (int i = 0; i < c->pn; i++) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->pev[2 * i], c->pev[2 * i + 1]) == hipSuccess) {
      c->prof_ms[c->pstage[i]] += ms;
      c->prof_calls[c->pstage[i]] += 1;
    }
  }
  c->pn = 0;
}

namespace {

enum {
  SEG_BOND = 1, SEG_ANGL = 2, SEG_TORS = 3, SEG_TORSRB = 4, SEG_14 = 5,
  SEG_PAIR = 10, SEG_EXCL = 11, SEG_PME = 12
};

// ---- host set-up mirrors of the reference's initialisation -----------------
void spline_coeff(int n, double h, const double *x, const double *y, double *b, double *c,
                  double *d) {
  std::vector<double> al(n + 1), l(n + 1), mu(n + 1), z(n + 1);
  for (int i = 2; i <= n - 1; i++)
    al[i] = 3.0 / h * ((y[i] - y[i - 1]) - (y[i - 1] - y[i - 2]));
  l[1] = 1.0;
  mu[1] = 0.0;
  z[1] = 0.0;
  for (int i = 2; i <= n - 1; i++) {
    l[i] = 2.0 * (x[i] - x[i - 2]) - h * mu[i - 1];
    mu[i] = h / l[i];
    z[i] = (al[i] - h * z[i - 1]) / l[i];
  }
  c[n - 1] = 0.0;
  for (int i = n - 1; i >= 1; i--) {
    c[i - 1] = z[i] - mu[i] * c[i];
    b[i - 1] = (y[i] - y[i - 1]) / h - h * (c[i] + 2.0 * c[i - 1]) / 3.0;
    d[i - 1] = (c[i] - c[i - 1]) / (3.0 * h);
  }
}

void build_splines(pgk_ctx *c, std::vector<double> &real8, std::vector<double> &excs8) {
  const int n = c->nspl;
  const double pie = 2.0 / sqrt(acos(-1.0)) * c->alpha;
  c->spl_int = (c->rrcut + 2.0) / (double)n;
  std::vector<double> x(n), tb[4], co[12];
  for (int k = 0; k < 4; k++) tb[k].assign(n, 0.0);
  for (int k = 0; k < 12; k++) co[k].assign(n, 0.0);
  for (int i = 0; i < n; i++) x[i] = (double)i * c->spl_int;
  tb[0][0] = tb[1][0] = tb[2][0] = tb[3][0] = 1.0;
  for (int i = 1; i < n; i++) {
    double r = x[i], inv = 1.0 / r, arg = c->alpha * r;
    tb[0][i] = inv * (pie * exp(-arg * arg) * inv + erfc(arg) * inv * inv);
    tb[1][i] = erfc(arg) * inv;
    tb[2][i] = inv * (-pie * exp(-arg * arg) * inv + (1.0 - erfc(arg)) * inv * inv);
    tb[3][i] = (1.0 - erfc(arg)) * inv;
  }
  for (int t = 0; t < 4; t++)
    spline_coeff(n, c->spl_int, x.data(), tb[t].data(), co[3 * t].data(), co[3 * t + 1].data(),
                 co[3 * t + 2].data());
  real8.assign((size_t)8 * n, 0.0);
  excs8.assign((size_t)8 * n, 0.0);
  for (int i = 0; i < n; i++) {
    for (int h = 0; h < 2; h++) {
      int t = h;       // 0 real (force), 1 realpot
      real8[8 * i + 4 * h + 0] = tb[t][i];
      real8[8 * i + 4 * h + 1] = co[3 * t][i];
      real8[8 * i + 4 * h + 2] = co[3 * t + 1][i];
      real8[8 * i + 4 * h + 3] = co[3 * t + 2][i];
      t = 2 + h;       // 2 excs (force), 3 excspot
      excs8[8 * i + 4 * h + 0] = tb[t][i];
      excs8[8 * i + 4 * h + 1] = co[3 * t][i];
      excs8[8 * i + 4 * h + 2] = co[3 * t + 1][i];
      excs8[8 * i + 4 * h + 3] = co[3 * t + 2][i];
    }
  }
}

void host_bspline(double w, int order, double *A, double *D) {   // 1-based views
  A[order] = 0.0;
  A[2] = w;
  A[1] = 1.0 - w;
  for (int k = 3; k <= order - 1; k++) {
    double div = 1.0 / (double)(k - 1);
    A[k] = div * w * A[k - 1];
    for (int j = 1; j <= k - 2; j++)
      A[k - j] = div * ((w + j) * A[k - j - 1] + (k - j - w) * A[k - j]);
    A[1] = div * (1 - w) * A[1];
  }
  D[1] = -A[1];
  for (int j = 2; j <= order; j++) D[j] = A[j - 1] - A[j];
  double div = 1.0 / (double)(order - 1);
  A[order] = div * w * A[order - 1];
  for (int j = 1; j <= order - 2; j++)
    A[order - j] = div * ((w + j) * A[order - j - 1] + (order - j - w) * A[order - j]);
  A[1] = div * (1 - w) * A[1];
}

void dftmod(std::vector<double> &mod, const std::vector<double> &arr, int n) {
  const double twopi = 2.0 * 3.14159265358979323846, tiny = 1.0e-7;
  mod.assign(n, 0.0);
  for (int k = 1; k <= n; k++) {
    double s1 = 0.0, s2 = 0.0;
    for (int j = 1; j <= n; j++) {
      double arg = twopi * (k - 1) * (j - 1) / n;
      s1 = s1 + arr[j - 1] * cos(arg);
      s2 = s2 + arr[j - 1] * sin(arg);
    }
    mod[k - 1] = s1 * s1 + s2 * s2;
  }
  for (int k = 1; k <= n; k++)
    if (mod[k - 1] < tiny) {
      double lo = (k >= 2) ? mod[k - 2] : mod[k - 1];
      double hi = (k < n) ? mod[k] : mod[k - 1];
      mod[k - 1] = 0.5 * (lo + hi);
    }
}

template <class T>
int upload(T **d, const std::vector<T> &h) {
  size_t n = h.size() ? h.size() : 1;
  if (*d) (void)hipFree(*d);
  *d = nullptr;
  HIP_CHECK(hipMalloc((void **)d, n * sizeof(T)));
  if (!h.empty()) HIP_CHECK(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return PGK_OK;
}

template <class T>
std::vector<T> copyv(const T *p, size_t n) {
  return (p && n) ? std::vector<T>(p, p + n) : std::vector<T>();
}

// ---- small kernels ---------------------------------------------------------
template <class R4>
__global__ void k_pack(int nslot, const int32_t *__restrict__ cl_atom,
                       const double *__restrict__ pos, const double *__restrict__ q,
                       const int32_t *__restrict__ type, R4 *__restrict__ xq,
                       int32_t *__restrict__ ct) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslot) return;
  int a = cl_atom[s];
  R4 v;
  if (a >= 0) {
    v.x = pos[3 * a];
    v.y = pos[3 * a + 1];
    v.z = pos[3 * a + 2];
    v.w = q[a];
    ct[s] = type[a] - 1;   // 0-based LJ type index
  } else {
    v.x = v.y = v.z = v.w = 0;
    ct[s] = 0;
  }
  xq[s] = v;
}

// deterministic segment sums: one block per (segment, field), fixed-order tree
__global__ __launch_bounds__(256) void k_reduce_segments(const double *__restrict__ red,
                                                         const int *__restrict__ seg_off,
                                                         const int *__restrict__ seg_n,
                                                         double *__restrict__ out) {
  __shared__ double s[256];
  const int sg = blockIdx.x / PGK_REC, f = blockIdx.x - sg * PGK_REC;
  const double *r = red + (size_t)seg_off[sg] * PGK_REC + f;
  const int n = seg_n[sg];
  double v = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) v += r[(size_t)k * PGK_REC];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[sg * PGK_REC + f] = s[0];
}

int set_device_ctx(pgk_ctx *c) {
  HIP_CHECK(hipSetDevice(c->device));
  return PGK_OK;
}

void sym6_to9(const double *v6, double *t9, double s) {
  // v6: xx yy zz xy xz yz  -> Fortran (a,n): idx a + 3n
  t9[0] += s * v6[0];
  t9[4] += s * v6[1];
  t9[8] += s * v6[2];
  t9[3] += s * v6[3];   // (x,y)
  t9[1] += s * v6[3];   // (y,x)
  t9[6] += s * v6[4];   // (x,z)
  t9[2] += s * v6[4];
  t9[7] += s * v6[5];   // (y,z)
  t9[5] += s * v6[5];
}

// One force call on the context's stream.  With hf (calforce_hf) the bonded
// kernels also scatter the per-atom heat-flux terms, the pair heat-flux pass
// runs over the list and the per-atom arrays are exported to the staging
// buffers before the final synchronisation.
int finalize_energies(pgk_ctx *c, double *ene, double *virt);

int run_calforce(pgk_ctx *c, int32_t mts_flag, const pgk_mts *m, int32_t preatom, bool split,
                 double *d_force_out, double *d_fvc, double *d_fvl, double *ene, double *virt,
                 bool hf = false, bool wait = true, bool energies = true) {
  if (!c->have_pos) {
    pgk_set_error("pgk_calforce: positions not set");
    return PGK_ESTATE;
  }
  const bool do_pair = (m->vdw == mts_flag);
  // the heat-flux pass over the listed pairs: with the pair forces, or on its
  // own (hf_pairs_only: the MD loop's heat-flux samples, forces already known)
  const bool hf_pairs = hf && (do_pair || c->hf_pairs_only);
  if ((do_pair || hf_pairs) && !c->have_list) {
    pgk_set_error("pgk_calforce: neighbour list not built");
    return PGK_ESTATE;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  hipStream_t st = c->stream;
  const int n = c->natom;
  const size_t ab = (c->precision == PGK_PREC_FP32) ? 4 : 8;
  // capacity of the record buffer: every launch appends its blocks
  int need = 16;
  if (do_pair)
    need += pgk_grid_blocks(std::max(c->ncl, c->nwi), 4) + pgk_grid_blocks(c->natom, 256) + 1;
  {
    int H = c->nfft[0] / 2 + 1;
    need += c->nfft[1] * H + 16;
  }
  need += pgk_grid_blocks((int64_t)c->nbond_act, 256) +
          pgk_grid_blocks((int64_t)c->nangl_act, 256) +
          pgk_grid_blocks((int64_t)c->tors_type.size(), 256) +
          pgk_grid_blocks((int64_t)c->rb_type.size(), 256) + pgk_grid_blocks(c->n14, 256) + 8;
  if (need > c->red_cap) {
    if (c->d_red) (void)hipFree(c->d_red);
    c->d_red = nullptr;
    HIP_CHECK(hipMalloc(&c->d_red, (size_t)need * PGK_REC * sizeof(double)));
    c->red_cap = need;
  }
  c->red_used = 0;
  int seg_off[16], seg_n[16], seg_kind[16], nseg = 0;

  if (c->acc_dirty) {   // first call, a new list size, or an interrupted call
    HIP_CHECK(hipMemsetAsync(c->d_fat, 0, (size_t)n * 3 * ab, st));
    HIP_CHECK(hipMemsetAsync(c->d_fat_c, 0, (size_t)n * 3 * ab, st));
    if (c->ncl_cap > 0) {
      size_t nsc = (size_t)c->ncl_cap * PGK_CL;
      HIP_CHECK(hipMemsetAsync(c->d_fcl, 0, nsc * 3 * ab, st));
      HIP_CHECK(hipMemsetAsync(c->d_fcl_c, 0, nsc * 3 * ab, st));
    }
    if (c->d_tile && c->tile_ntile > 0)   // spreading bin counts
      HIP_CHECK(hipMemsetAsync(c->d_tile, 0, sizeof(int32_t) * c->tile_ntile, st));
  }
  c->acc_dirty = true;
  const bool do_ewk = (m->ewk == mts_flag);
  if (do_ewk) {   // B-spline weights and tile ranks; packs the pair data too
    pgk_prof_begin(c, PGK_PROF_OTHER);
    if ((rc = pgk_pme_prep(c, do_pair))) return rc;
    pgk_prof_end(c, PGK_PROF_OTHER);
  }
  // The reciprocal chain (spread, FFT, gather) and the pair kernel read the
  // same positions and write different accumulators: with both due, the
  // chain goes to the second stream and the pair and bonded kernels run
  // beside it (timing any stage but the pair kernel keeps them one after
  // the other)
  const bool timed_other = c->prof && (c->prof_mask & ~(1u << PGK_PROF_PAIRS)) != 0;
  const bool overlap = do_pair && do_ewk && !hf && !timed_other && c->pme_overlap;
  // beside the chain the excluded-pair correction goes to the shorter stream:
  // the pair kernel's where the chain outlasts the pair and bonded kernels
  // (small systems in the chunked form; the FP64 form, whose chain runs in
  // FP64: 0.376 -> 0.368 ms per step on the 100k water box), the chain's
  // otherwise (mixed and FP32 on large systems: 0.291 ms there against 0.296)
  const bool excl_main =
      overlap && (pgk_pair_chunked(c) || c->precision == PGK_PREC_FP64);
  if (overlap) {
    HIP_CHECK(hipEventRecord(c->ev_fork, st));
    HIP_CHECK(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
    c->stream = c->stream2;
    rc = pgk_pme_launch(c, split, &seg_off[nseg], &seg_n[nseg]);
    if (!rc) {
      seg_kind[nseg++] = SEG_PME;
      if (!excl_main) {
        rc = pgk_excl_launch(c, split, &seg_off[nseg], &seg_n[nseg]);
        if (!rc && seg_n[nseg] > 0) seg_kind[nseg++] = SEG_EXCL;
      }
    }
    c->stream = st;
    if (rc) return rc;
    c->pme_grid_ready = true;
    HIP_CHECK(hipEventRecord(c->ev_join, c->stream2));
  }
  if (do_pair) {
    size_t ns = (size_t)c->ncl * PGK_CL;
    if (!do_ewk) {
      if (c->precision == PGK_PREC_FP64)
        k_pack<double4><<<pgk_grid_blocks(ns, 256), 256, 0, st>>>(
            (int)ns, c->d_cl_atom, c->d_pos, c->d_q, c->d_type, (double4 *)c->d_xq, c->d_ct);
      else
        k_pack<float4><<<pgk_grid_blocks(ns, 256), 256, 0, st>>>(
            (int)ns, c->d_cl_atom, c->d_pos, c->d_q, c->d_type, (float4 *)c->d_xq, c->d_ct);
      HIP_CHECK(hipGetLastError());
    }
    pgk_prof_begin(c, PGK_PROF_PAIRS);
    // beside the chain, a cap on the pair blocks per CU (the chunked form's
    // few hundred blocks fill the device in one round at three per CU)
    c->pair_cap_blocks =
        !overlap ? 0
                 : (pgk_pair_chunked(c) && c->pair_cap_default > 0 ? c->pair_cap_chunked
                                                                  : c->pair_cap_default);
    rc = pgk_pairs_launch(c, split, energies, &seg_off[nseg], &seg_n[nseg]);
    c->pair_cap_blocks = 0;
    if (rc) return rc;
    pgk_prof_end(c, PGK_PROF_PAIRS);
    seg_kind[nseg++] = SEG_PAIR;
  }
  if (do_ewk && !overlap) {
    if ((rc = pgk_pme_launch(c, split, &seg_off[nseg], &seg_n[nseg]))) return rc;
    seg_kind[nseg++] = SEG_PME;
    c->pme_grid_ready = true;
  }
  pgk_prof_begin(c, PGK_PROF_BONDED);
  if ((rc = pgk_bonded_launch2(c, m, mts_flag, seg_off, seg_n, seg_kind, &nseg, hf))) return rc;
  pgk_prof_end(c, PGK_PROF_BONDED);
  if (overlap) {
    if (excl_main) {
      if ((rc = pgk_excl_launch(c, split, &seg_off[nseg], &seg_n[nseg]))) return rc;
      if (seg_n[nseg] > 0) seg_kind[nseg++] = SEG_EXCL;
    }
    HIP_CHECK(hipStreamWaitEvent(st, c->ev_join, 0));
  } else if (do_pair) {
    pgk_prof_begin(c, PGK_PROF_EXCL);
    if ((rc = pgk_excl_launch(c, split, &seg_off[nseg], &seg_n[nseg]))) return rc;
    pgk_prof_end(c, PGK_PROF_EXCL);
    if (seg_n[nseg] > 0) seg_kind[nseg++] = SEG_EXCL;
  }
  if (hf_pairs && (rc = pgk_hf_launch_pairs(c, true))) return rc;
  // forces: accumulators + excluded-pair correction -> FP64 output (a
  // heat-flux sample of the MD loop adds no forces and has none to sum)
  double *fout = d_force_out ? d_force_out : c->d_force;
  if (!c->hf_pairs_only &&
      (rc = pgk_combine_launch(c, do_pair, split, fout, split ? d_fvc : nullptr,
                               split ? d_fvl : nullptr)))
    return rc;
  // energies: fixed-order segment sums into d_sum, copied to pinned h_sum
  // (a dynamics step that samples nothing skips both: nothing reads them)
  if (nseg > 0 && energies) {
    // table slots: the MD loop's heat-flux pass and RESPA's short-level calls
    // alternate with the step's own call without re-uploading any table
    const int slot = c->hf_pairs_only ? 1 : (mts_flag == PGK_MTS_SHORT ? 2 : 0);
    int *dseg = reinterpret_cast<int *>(c->d_sum + 512 + 64 * slot);
    int *hseg = reinterpret_cast<int *>(c->h_sum + 512 + 64 * slot);   // pinned staging
    int tab[32] = {};
    for (int k = 0; k < nseg; k++) {
      tab[k] = seg_off[k];
      tab[16 + k] = seg_n[k];
    }
    if (c->seg_up_n[slot] != nseg || memcmp(tab, c->seg_up[slot], sizeof(tab)) != 0) {
      // an earlier queued upload may still read the staging area
      HIP_CHECK(hipStreamSynchronize(st));
      memcpy(hseg, tab, sizeof(tab));
      HIP_CHECK(hipMemcpyAsync(dseg, hseg, 32 * sizeof(int), hipMemcpyHostToDevice, st));
      memcpy(c->seg_up[slot], tab, sizeof(tab));
      c->seg_up_n[slot] = nseg;
    }
    k_reduce_segments<<<nseg * PGK_REC, 256, 0, st>>>(c->d_red, dseg, dseg + 16, c->d_sum);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(c->h_sum, c->d_sum, sizeof(double) * nseg * PGK_REC,
                             hipMemcpyDeviceToHost, st));
  }
  if (hf && c->hf_export && (rc = pgk_hf_export(c))) return rc;
  c->last_nseg = nseg;
  for (int k = 0; k < nseg; k++) c->last_seg_kind[k] = seg_kind[k];
  c->last_mts_flag = mts_flag;
  c->last_mts = *m;
  c->last_pending = true;
  c->acc_dirty = false;   // the queued kernels consume and clear every accumulator
  if (!wait) return PGK_OK;
  return finalize_energies(c, ene, virt);
}

// Wait for the last call and turn its segment sums into the energy record
// and virial tensors (calforce.F90:1018-1068 for the net-charge terms and
// the totals).  ene is inout: energies of terms not evaluated keep their
// values.
int finalize_energies(pgk_ctx *c, double *ene, double *virt) {
  HIP_CHECK(hipStreamSynchronize(c->stream));
  pgk_prof_collect(c);
  if (!c->last_pending) {
    pgk_set_error("no calforce call to read energies from");
    return PGK_ESTATE;
  }
  const int nseg = c->last_nseg;
  const int *seg_kind = c->last_seg_kind;
  const pgk_mts *m = &c->last_mts;
  const int32_t mts_flag = c->last_mts_flag;
  const bool do_pair = (m->vdw == mts_flag);
  double segsum[16][PGK_REC];
  for (int k = 0; k < nseg; k++) memcpy(segsum[k], c->h_sum + k * PGK_REC, sizeof(segsum[k]));

  double E[PGK_NENE];
  double T[PGK_NTENS][9];
  memcpy(E, ene, sizeof(E));   // energies not evaluated keep the caller's values
  memset(T, 0, sizeof(T));
  E[PGK_V_COUL] = E[PGK_V_LJ] = 0.0;
  E[PGK_V_BOND] = E[PGK_V_ANGL] = E[PGK_V_TORS] = E[PGK_V_14] = 0.0;
  bool tors_done = false;
  // the pair record sets the Coulomb and LJ sums that the excluded-pair
  // record then corrects, whichever stream queued them first
  int korder[16], nk = 0;
  for (int k = 0; k < nseg; k++)
    if (seg_kind[k] == SEG_PAIR) korder[nk++] = k;
  for (int k = 0; k < nseg; k++)
    if (seg_kind[k] != SEG_PAIR) korder[nk++] = k;
  for (int kk = 0; kk < nseg; kk++) {
    const int k = korder[kk];
    const double *s = segsum[k];
    switch (seg_kind[k]) {
      case SEG_PAIR:
        E[PGK_E_ELC] = s[0];
        E[PGK_E_VDW] = s[1];
        E[PGK_V_COUL] += s[0];
        sym6_to9(s + 2, T[PGK_T_COUL], 1.0);
        sym6_to9(s + 8, T[PGK_T_LJ], 1.0);
        E[PGK_V_LJ] = s[8] + s[9] + s[10];
        break;
      case SEG_EXCL:
        E[PGK_E_ELC] += s[0];
        E[PGK_V_COUL] += s[0];
        sym6_to9(s + 1, T[PGK_T_COUL], 1.0);
        break;
      case SEG_PME: {
        E[PGK_E_EWK] = 0.5 * s[0];
        double vsca = 0.5 * s[1];
        E[PGK_V_COUL] += E[PGK_E_EWK] + vsca;
        sym6_to9(s + 2, T[PGK_T_COUL], -0.5);
        break;
      }
      case SEG_BOND:
        E[PGK_E_BOND] = s[0];
        E[PGK_V_BOND] += s[1];
        for (int q = 0; q < 9; q++) T[PGK_T_BOND][q] += s[2 + q];
        break;
      case SEG_ANGL:
        E[PGK_E_ANGL] = s[0];
        E[PGK_V_ANGL] += s[1];
        for (int q = 0; q < 9; q++) T[PGK_T_ANGL][q] += s[2 + q];
        break;
      case SEG_TORS:
      case SEG_TORSRB:
        E[seg_kind[k] == SEG_TORS ? PGK_E_TORS : PGK_E_TORSRB] = s[0];
        E[PGK_V_TORS] += s[1];
        for (int q = 0; q < 9; q++) T[PGK_T_TORS][q] += s[2 + q];
        tors_done = true;
        break;
      case SEG_14:
        E[PGK_E_ELC14] = s[0];
        E[PGK_E_VDW14] = s[1];
        E[PGK_V_14] += s[2];
        for (int q = 0; q < 9; q++) T[PGK_T_14][q] += s[3 + q];
        break;
    }
  }
  (void)tors_done;
  // terms whose MTS level matches but that have no entries still report zero
  if (m->bond == mts_flag && c->nbond_act == 0) E[PGK_E_BOND] = 0.0;
  if (m->angl == mts_flag && c->nangl_act == 0) E[PGK_E_ANGL] = 0.0;
  if (m->tors == mts_flag && c->tors_type.empty()) E[PGK_E_TORS] = 0.0;
  if (m->torsrb == mts_flag && c->rb_type.empty()) E[PGK_E_TORSRB] = 0.0;
  if (m->nonb14 == mts_flag && c->n14 == 0) E[PGK_E_ELC14] = E[PGK_E_VDW14] = 0.0;
  if (do_pair && c->nexclpair == 0) { /* pot_elc already set from the pairs */ }
  E[PGK_E_EWC] = c->pot_ewc;
  if (mts_flag == PGK_MTS_LONG) {   // calforce.F90:1018-1044
    double pi = acos(-1.0);
    double a1 = -pi / (c->alpha * c->alpha * c->box[0] * c->box[1] * c->box[2]);
    E[PGK_E_EWNQ] = 0.5 * a1 * c->netchrgsq;
    if (!c->ifnetq) {
      E[PGK_V_COUL] += a1 * c->netchrgsq;
      for (int d = 0; d < 3; d++) T[PGK_T_COUL][4 * d] += 0.5 * a1 * c->netchrgsq;
    } else {
      E[PGK_V_COUL] -= 0.5 * a1 * c->netchrgsq;
    }
  }
  E[PGK_E_NONBON] = E[PGK_E_VDW] + E[PGK_E_ELC] + E[PGK_E_EWK] + E[PGK_E_EWC] + E[PGK_E_EWNQ] +
                    E[PGK_E_VDW14] + E[PGK_E_ELC14];
  E[PGK_E_TOT] = E[PGK_E_BOND] + E[PGK_E_ANGL] + E[PGK_E_TORS] + E[PGK_E_TORSRB] +
                 E[PGK_E_NONBON];
  memcpy(ene, E, sizeof(E));
  if (virt) memcpy(virt, T, sizeof(T));
  return PGK_OK;
}

}  // namespace

// =============================================================================
extern "C" {

int32_t pgk_abi_version(void) { return PGK_ABI_VERSION; }
const char *pgk_last_error(void) { return g_err.c_str(); }

// Largest double T with sqrt(T) <= rc in IEEE arithmetic: since the
// correctly rounded sqrt is monotone, r2 <= T decides sqrt(r2) <= rc exactly.
static double sqrt_threshold(double rc) {
  double t = rc * rc;
  while (std::sqrt(t) > rc) t = std::nextafter(t, 0.0);
  while (std::sqrt(std::nextafter(t, HUGE_VAL)) <= rc) t = std::nextafter(t, HUGE_VAL);
  return t;
}

int pgk_create(const pgk_system_desc *d, int32_t precision, pgk_ctx **out) {
  if (!d || !out) {
    pgk_set_error("pgk_create: null argument");
    return PGK_EINVAL;
  }
  *out = nullptr;
  if (d->natom <= 1 || d->ntype <= 0 || !d->charge || !d->atype || !d->molid || !d->lj_eps ||
      !d->lj_rmin || !d->inttyp || !d->excl_index || !d->box) {
    pgk_set_error("pgk_create: incomplete system description");
    return PGK_EINVAL;
  }
  if (precision < PGK_PREC_FP64 || precision > PGK_PREC_FP32) {
    pgk_set_error("pgk_create: unknown precision");
    return PGK_EINVAL;
  }
  if (d->nspltbl < 4 || d->nfft1 < 4 || d->nfft2 < 4 || d->nfft3 < 4) {
    pgk_set_error("pgk_create: spline table or SPME grid too small");
    return PGK_EINVAL;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    pgk_set_error("pgk_create: no HIP device available");
    return PGK_EDEVICE;
  }
  const int n = d->natom, nt = d->ntype;
  for (int i = 0; i < n; i++)
    if (d->atype[i] < 1 || d->atype[i] > nt) {
      pgk_set_error("pgk_create: atom type out of range");
      return PGK_EINVAL;
    }
  int dev = 0;
  (void)hipGetDevice(&dev);
  {   // the allocation-free entry points' scratch address on this device
    const int rc0 = pgk_noalloc_init_device(dev);
    if (rc0) return rc0;
  }
  pgk_ctx *c = new pgk_ctx();
  c->precision = precision;
  c->device = dev;
  c->natom = n;
  c->ntype = nt;
  c->charge = copyv(d->charge, n);
  c->atype = copyv(d->atype, n);
  c->molid = copyv(d->molid, n);
  c->lj_eps = copyv(d->lj_eps, (size_t)nt * nt);
  c->lj_rmin = copyv(d->lj_rmin, (size_t)nt * nt);
  c->inttyp = copyv(d->inttyp, (size_t)nt * nt);
  c->excl_index = copyv(d->excl_index, (size_t)n + 1);
  c->excl_list = copyv(d->excl_list, (size_t)d->nexcl);
  c->bond_ij = copyv(d->bond_ij, (size_t)2 * d->nbond);
  c->bond_type = copyv(d->bond_type, (size_t)d->nbond);
  c->bond_k = copyv(d->bond_k, (size_t)d->nbondtyp);
  c->bond_r0 = copyv(d->bond_r0, (size_t)d->nbondtyp);
  c->angl_ijk = copyv(d->angl_ijk, (size_t)3 * d->nangl);
  c->angl_type = copyv(d->angl_type, (size_t)d->nangl);
  c->angl_k = copyv(d->angl_k, (size_t)d->nangltyp);
  c->angl_t0 = copyv(d->angl_theta0, (size_t)d->nangltyp);
  c->tors_ijkl = copyv(d->tors_ijkl, (size_t)4 * d->ntors);
  c->tors_type = copyv(d->tors_type, (size_t)d->ntors);
  c->tors_ring = copyv(d->tors_ring14, (size_t)d->ntors);
  c->tors_divfac = copyv(d->tors_divfac, (size_t)d->ntorstyp);
  c->tors_barhig = copyv(d->tors_barhig, (size_t)d->ntorstyp);
  c->tors_phsang = copyv(d->tors_phsang, (size_t)d->ntorstyp);
  c->tors_period = copyv(d->tors_period, (size_t)d->ntorstyp);
  c->tors_dvdw = copyv(d->tors_divfac_vdw, (size_t)d->ntorstyp);
  c->tors_delc = copyv(d->tors_divfac_elc, (size_t)d->ntorstyp);
  c->tors_if14 = copyv(d->tors_if14, (size_t)d->ntorstyp);
  c->rb_ijkl = copyv(d->torsrb_ijkl, (size_t)4 * d->ntorsrb);
  c->rb_type = copyv(d->torsrb_type, (size_t)d->ntorsrb);
  c->rb_ring = copyv(d->torsrb_ring14, (size_t)d->ntorsrb);
  c->rb_c = copyv(d->torsrb_c, (size_t)6 * d->ntorsrbtyp);
  c->rb_dvdw = copyv(d->torsrb_divfac_vdw, (size_t)d->ntorsrbtyp);
  c->rb_delc = copyv(d->torsrb_divfac_elc, (size_t)d->ntorsrbtyp);
  c->rb_if14 = copyv(d->torsrb_if14, (size_t)d->ntorsrbtyp);
  c->alpha = d->alpha;
  c->rrcut = d->rrcut;
  c->rrsq_cut = sqrt_threshold(d->rrcut);
  c->rcut = d->rcut;
  c->nspl = d->nspltbl;
  c->nfft[0] = d->nfft1;
  c->nfft[1] = d->nfft2;
  c->nfft[2] = d->nfft3;
  c->order = d->pme_order;
  c->ifnetq = d->ifnetqcorrp;
  for (int k = 0; k < 3; k++) c->box[k] = d->box[k];
  // validate topology indices
  auto bad = [&](const std::vector<int32_t> &v) {
    for (int32_t x : v)
      if (x < 1 || x > n) return true;
    return false;
  };
  auto badt = [&](const std::vector<int32_t> &v, int ntyp) {
    for (int32_t x : v)
      if (x < 1 || x > ntyp) return true;
    return false;
  };
  bool bad_csr = (int)c->excl_index.size() != n + 1 || c->excl_index[0] != 1 ||
                 c->excl_index[n] - 1 != d->nexcl;
  for (int i = 0; i < n && !bad_csr; i++)
    if (c->excl_index[i + 1] < c->excl_index[i]) bad_csr = true;
  for (int32_t x : c->excl_list)
    if (x < 0 || x > n) bad_csr = true;
  if (bad(c->bond_ij) || bad(c->angl_ijk) || bad(c->tors_ijkl) || bad(c->rb_ijkl) ||
      badt(c->bond_type, d->nbondtyp) || badt(c->angl_type, d->nangltyp) ||
      badt(c->tors_type, d->ntorstyp) || badt(c->rb_type, d->ntorsrbtyp) || bad_csr ||
      (int)c->tors_ring.size() != d->ntors || (int)c->rb_ring.size() != d->ntorsrb) {
    pgk_set_error("pgk_create: topology index, type or exclusion CSR out of range");
    delete c;
    return PGK_EINVAL;
  }
  for (int i = 0; i < n; i++)
    if (c->molid[i] < 1) {
      pgk_set_error("pgk_create: molecule index out of range");
      delete c;
      return PGK_EINVAL;
    }
  int rc = pgk_pme_prepare(c);
  if (rc) {
    delete c;
    return rc;
  }
  // netchrgsq, self energy, moduli, rkcut
  double qs = 0.0, q2 = 0.0;
  for (int i = 0; i < n; i++) {
    qs = qs + c->charge[i];
    q2 = q2 + c->charge[i] * c->charge[i];
  }
  c->netchrgsq = qs * qs;
  c->pot_ewc = -1.0 / sqrt(acos(-1.0)) * c->alpha * q2;
  {
    double A[PGK_MAXORD + 2], D[PGK_MAXORD + 2];
    host_bspline(0.0, c->order, A, D);
    int maxn = std::max(c->nfft[0], std::max(c->nfft[1], c->nfft[2]));
    std::vector<double> arr(maxn + PGK_MAXORD + 2, 0.0);
    for (int i = 2; i <= c->order + 1; i++) arr[i - 1] = A[i - 1];
    for (int dd = 0; dd < 3; dd++) dftmod(c->bsp_mod[dd], arr, c->nfft[dd]);
    double m1 = 0.5 * (c->nfft[0] / d->box[0]);
    double rk = m1;
    double m2 = 0.5 * (c->nfft[1] / d->box[1]);
    if (m2 < m1) rk = m2;
    double m3 = 0.5 * (c->nfft[2] / d->box[2]);
    if (m3 < m2) rk = m3;
    c->rkcut = 0.9 * rk;
  }
  hipError_t he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (he == hipSuccess) he = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  if (he == hipSuccess &&
      hipDeviceGetAttribute(&c->lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                            c->device) != hipSuccess)
    c->lds_per_cu = 0;
  if (he != hipSuccess) {
    pgk_set_error(std::string("pgk_create: stream: ") + hipGetErrorString(he));
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PGK_EDEVICE;
  }
  // ---- device uploads ----
  std::vector<double> real8, excs8;
  build_splines(c, real8, excs8);
  std::vector<float> realf(real8.begin(), real8.end());
  std::vector<double> ljp64((size_t)2 * nt * nt);
  std::vector<float> ljp32((size_t)2 * nt * nt);
  for (size_t k = 0; k < (size_t)nt * nt; k++) {
    // a type pair that is not INTTYPE_VDW contributes no LJ (ewaldr.F90:260):
    // fold that into a zero well depth
    ljp64[2 * k] = (c->inttyp[k] == 1) ? c->lj_eps[k] : 0.0;
    ljp64[2 * k + 1] = c->lj_rmin[k] * c->lj_rmin[k];
    ljp32[2 * k] = (float)ljp64[2 * k];
    ljp32[2 * k + 1] = (float)ljp64[2 * k + 1];
  }
  std::vector<int32_t> exclpair;
  for (int i = 0; i < n; i++)
    for (int k = c->excl_index[i] - 1; k < c->excl_index[i + 1] - 1; k++) {
      int j = c->excl_list[k];
      if (j > 0) {
        exclpair.push_back(i);
        exclpair.push_back(j - 1);
      }
    }
  c->nexclpair = (int)(exclpair.size() / 2);
  std::vector<int32_t> xs_start(n + 1, 0), xs_j(exclpair.size());
  for (size_t k = 0; k < exclpair.size(); k++) xs_start[exclpair[k] + 1]++;
  for (int i = 0; i < n; i++) xs_start[i + 1] += xs_start[i];
  {
    std::vector<int32_t> fill(xs_start.begin(), xs_start.end() - 1);
    for (size_t k = 0; k < exclpair.size(); k += 2) {
      const int i = exclpair[k], j = exclpair[k + 1];
      xs_j[fill[i]++] = j;
      xs_j[fill[j]++] = i;
    }
  }
  // the list search removes a pair i<j only when j is on i's own list
  // (mklist2a), so its symmetric copy keeps just the entries above the row
  std::vector<int32_t> xl_start(n + 1, 0), xl_j;
  {
    std::vector<int32_t> up;
    for (size_t k = 0; k < exclpair.size(); k += 2)
      if (exclpair[k + 1] > exclpair[k]) {
        up.push_back(exclpair[k]);
        up.push_back(exclpair[k + 1]);
      }
    xl_j.resize(up.size());
    for (size_t k = 0; k < up.size(); k++) xl_start[up[k] + 1]++;
    for (int i = 0; i < n; i++) xl_start[i + 1] += xl_start[i];
    std::vector<int32_t> fill(xl_start.begin(), xl_start.end() - 1);
    for (size_t k = 0; k < up.size(); k += 2) {
      xl_j[fill[up[k]]++] = up[k + 1];
      xl_j[fill[up[k + 1]]++] = up[k];
    }
  }
  // bonded: 0-based indices, parameters packed
  std::vector<int32_t> hb, ha, ht, hr, hp14;
  std::vector<double> par, p14f;
  c->off_bk = (int)par.size();
  par.insert(par.end(), c->bond_k.begin(), c->bond_k.end());
  c->off_br = (int)par.size();
  par.insert(par.end(), c->bond_r0.begin(), c->bond_r0.end());
  c->off_ak = (int)par.size();
  par.insert(par.end(), c->angl_k.begin(), c->angl_k.end());
  c->off_at = (int)par.size();
  par.insert(par.end(), c->angl_t0.begin(), c->angl_t0.end());
  c->off_tdiv = (int)par.size();
  for (int32_t x : c->tors_divfac) par.push_back((double)x);
  c->off_tb = (int)par.size();
  par.insert(par.end(), c->tors_barhig.begin(), c->tors_barhig.end());
  c->off_tp = (int)par.size();
  par.insert(par.end(), c->tors_phsang.begin(), c->tors_phsang.end());
  c->off_tper = (int)par.size();
  par.insert(par.end(), c->tors_period.begin(), c->tors_period.end());
  c->off_rbc = (int)par.size();
  par.insert(par.end(), c->rb_c.begin(), c->rb_c.end());
  // terms the reference skips for a vanishing force constant (|K| < 1e-16
  // for bonds, < 1e-32 for angles) are left out of the device lists
  for (size_t k = 0; k < c->bond_type.size(); k++) {
    if (fabs(c->bond_k[c->bond_type[k] - 1]) < 1.0e-16) continue;
    hb.push_back(c->bond_ij[2 * k] - 1);
    hb.push_back(c->bond_ij[2 * k + 1] - 1);
    hb.push_back(c->bond_type[k] - 1);
  }
  for (size_t k = 0; k < c->angl_type.size(); k++) {
    if (fabs(c->angl_k[c->angl_type[k] - 1]) < 1.0e-32) continue;
    for (int q = 0; q < 3; q++) ha.push_back(c->angl_ijk[3 * k + q] - 1);
    ha.push_back(c->angl_type[k] - 1);
  }
  c->nbond_act = (int)(hb.size() / 3);
  c->nangl_act = (int)(ha.size() / 4);
  for (size_t k = 0; k < c->tors_type.size(); k++) {
    int np = (int)lround(fabs(c->tors_period[c->tors_type[k] - 1]));
    if (np < 1 || np > 6) {
      pgk_set_error("pgk_create: torsion periodicity not supported (caltors.F90:414-435)");
      (void)hipStreamDestroy(c->stream);
      delete c;
      return PGK_EINVAL;
    }
    for (int q = 0; q < 4; q++) ht.push_back(c->tors_ijkl[4 * k + q] - 1);
    ht.push_back(c->tors_type[k] - 1);
  }
  for (size_t k = 0; k < c->rb_type.size(); k++) {
    for (int q = 0; q < 4; q++) hr.push_back(c->rb_ijkl[4 * k + q] - 1);
    hr.push_back(c->rb_type[k] - 1);
  }
  // 1-4 pairs: periodic torsions first, then RB (calnon14.F90:274-509)
  for (size_t k = 0; k < c->tors_type.size(); k++) {
    int t = c->tors_type[k] - 1;
    if (!c->tors_if14[t]) continue;
    int i = c->tors_ijkl[4 * k], l = c->tors_ijkl[4 * k + 3];
    if (i == l) continue;
    hp14.push_back(i - 1);
    hp14.push_back(l - 1);
    p14f.push_back(1.0 / c->tors_dvdw[t] * c->tors_ring[k]);
    p14f.push_back(1.0 / c->tors_delc[t] * c->tors_ring[k]);
  }
  for (size_t k = 0; k < c->rb_type.size(); k++) {
    int t = c->rb_type[k] - 1;
    if (!c->rb_if14[t]) continue;
    int i = c->rb_ijkl[4 * k], l = c->rb_ijkl[4 * k + 3];
    if (i == l) continue;
    hp14.push_back(i - 1);
    hp14.push_back(l - 1);
    p14f.push_back(1.0 / c->rb_dvdw[t] * c->rb_ring[k]);
    p14f.push_back(1.0 / c->rb_delc[t] * c->rb_ring[k]);
  }
  c->n14 = (int)(hp14.size() / 2);
  // twiddles exp(-2 pi i k/n)
  std::vector<double> tw;
  std::vector<float> twf;
  for (int dd = 0; dd < 3; dd++) {
    c->tw_off[dd] = (int)(tw.size() / 2);
    int nn = c->nfft[dd];
    for (int k = 0; k < nn; k++) {
      double a = 2.0 * 3.14159265358979323846 * (double)k / (double)nn;
      tw.push_back(cos(a));
      tw.push_back(-sin(a));
    }
  }
  twf.assign(tw.begin(), tw.end());
  std::vector<double> bsp;
  for (int dd = 0; dd < 3; dd++) bsp.insert(bsp.end(), c->bsp_mod[dd].begin(), c->bsp_mod[dd].end());

  if ((rc = upload(&c->d_q, c->charge)) || (rc = upload(&c->d_type, c->atype)) ||
      (rc = upload(&c->d_molid, c->molid)) || (rc = upload(&c->d_eps, c->lj_eps)) ||
      (rc = upload(&c->d_rmin, c->lj_rmin)) || (rc = upload(&c->d_inttyp, c->inttyp)) ||
      (rc = upload(&c->d_excl_index, c->excl_index)) ||
      (rc = upload(&c->d_excl_list, c->excl_list)) || (rc = upload(&c->d_exclpair, exclpair)) ||
      (rc = upload(&c->d_xs_start, xs_start)) || (rc = upload(&c->d_xs_j, xs_j)) ||
      (rc = upload(&c->d_xl_start, xl_start)) || (rc = upload(&c->d_xl_j, xl_j)) ||
      (rc = upload(&c->d_spl, real8)) || (rc = upload(&c->d_splx, excs8)) ||
      (rc = upload(&c->d_splf, realf)) || (rc = upload(&c->d_ljp64, ljp64)) ||
      (rc = upload(&c->d_ljp32, ljp32)) || (rc = upload(&c->d_bsp, bsp)) ||
      (rc = upload(&c->d_bond, hb)) || (rc = upload(&c->d_angl, ha)) ||
      (rc = upload(&c->d_tors, ht)) || (rc = upload(&c->d_rb, hr)) ||
      (rc = upload(&c->d_p14, hp14)) || (rc = upload(&c->d_p14f, p14f)) ||
      (rc = upload(&c->d_bpar, par)) || (rc = upload(&c->d_tw, tw)) ||
      (rc = upload(&c->d_twf, twf))) {
    pgk_destroy(c);
    return rc;
  }
  size_t ab = (precision == PGK_PREC_FP32) ? 4 : 8;
  size_t G = (size_t)c->nfft[0] * c->nfft[1] * c->nfft[2];
  size_t Hc = (size_t)(c->nfft[0] / 2 + 1) * c->nfft[1] * c->nfft[2];
  size_t csz = (precision == PGK_PREC_FP64) ? 16 : 8;
  he = hipMalloc(&c->d_pos, sizeof(double) * 3 * n);
  if (he == hipSuccess) he = hipMalloc(&c->d_fat, ab * 3 * n);
  if (he == hipSuccess) he = hipMalloc(&c->d_fat_c, ab * 3 * n);
  if (he == hipSuccess) he = hipMalloc((void **)&c->d_force, sizeof(double) * 3 * n);
  if (he == hipSuccess) he = hipMalloc((void **)&c->d_fvc, sizeof(double) * 3 * n);
  if (he == hipSuccess) he = hipMalloc((void **)&c->d_fvl, sizeof(double) * 3 * n);
  if (he == hipSuccess) he = hipMalloc(&c->d_grid, G * 8);
  if (he == hipSuccess) he = hipMalloc(&c->d_pgrid, G * ab);
  if (he == hipSuccess) he = hipMalloc(&c->d_cgrid, Hc * csz);
  if (he == hipSuccess) he = hipMalloc((void **)&c->d_sum, sizeof(double) * 1024);
  if (he == hipSuccess) he = hipHostMalloc((void **)&c->h_sum, sizeof(double) * 1024);
  if (he != hipSuccess) {
    pgk_set_error(std::string("pgk_create: device allocation failed: ") + hipGetErrorString(he));
    pgk_destroy(c);
    return PGK_ENOMEM;
  }
  *out = c;
  return PGK_OK;
}

int pgk_destroy(pgk_ctx *c) {
  if (!c) return PGK_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void *ptrs[] = {c->d_pos, c->d_q, c->d_type, c->d_molid, c->d_eps, c->d_rmin, c->d_inttyp,
                  c->d_excl_index, c->d_excl_list, c->d_exclpair, c->d_xs_start, c->d_xs_j,
                  c->d_xl_start, c->d_xl_j,
                  c->d_spl, c->d_splx,
                  c->d_splf, c->d_ljp64, c->d_ljp32, c->d_bsp, c->d_bond, c->d_angl, c->d_tors,
                  c->d_rb, c->d_p14, c->d_p14f, c->d_bpar, c->d_bin_count, c->d_bin_start,
                  c->d_atom_bin, c->d_atom_slot, c->d_sorted, c->d_col_ncl, c->d_col_clstart,
                  c->d_cl_atom, c->d_cl_col, c->d_slot_of, c->d_bbox, c->d_posc, c->d_tmp_cj,
                  c->d_list_meta,
                  c->d_tmp_mask, c->d_wcount, c->d_wstart, c->d_witem, c->d_ci_count, c->d_ci_start, c->d_cj, c->d_mask,
                  c->d_overflow, c->d_scan_tmp, c->d_xq, c->d_ct, c->d_fcl, c->d_fat,
                  c->d_fcl_c, c->d_fat_c, c->d_force, c->d_fvc, c->d_fvl, c->d_red, c->d_sum,
                  c->d_grid, c->d_pgrid, c->d_cgrid, c->d_tw, c->d_twf, c->d_theta, c->d_tile,
                  c->d_hf_z, c->d_hf_typ, c->d_hf_com, c->d_hf_pe, c->d_hf_w,
                  c->d_hf_pe_out, c->d_hf_w_out, c->d_md_gstart, c->d_md_cij, c->d_md_fail,
                  c->d_md_small,
                  c->d_md_d2, c->d_md_minv, c->d_md_vel, c->d_md_pos_old, c->d_md_free, c->d_md_ke,
                  c->d_md_kenhc,
                  c->d_md_nhc, c->d_lt_ustart, c->d_lt_uatom, c->d_lt_ucom, c->d_lt_reg,
                  c->d_lt_cnt, c->d_lt_mass, c->d_lt_frand, c->d_lt_det, c->d_lt_sreg, c->d_md_vstrm,
                  c->d_hf_vstrm, c->d_hj_vpec, c->d_vgrid,
                  c->d_vir6, c->d_hfw_xq, c->d_hfw_cen, c->d_hfw_mol, c->d_hfw_v, c->d_tl_start,
                  c->d_tl_src, c->d_tl_mask, c->d_hj_vel, c->d_hj_mass, c->d_hj_molevel,
                  c->d_hj_cz,
                  c->d_hj_part, c->d_hj_out, c->d_hj_mol_start, c->d_hj_mol_atoms,
                  c->d_hj_jcol, c->d_hj_jpart, c->d_md_hfseries, c->d_md_hfmass,
                  c->d_hj_fscratch, c->d_md_fshort};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->h_sum) (void)hipHostFree(c->h_sum);
  if (c->h_list_meta) (void)hipHostFree(c->h_list_meta);
  for (hipEvent_t e : c->pev)
    if (e) (void)hipEventDestroy(e);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PGK_OK;
}

int pgk_set_box(pgk_ctx *c, const double *box) {
  if (!c || !box || !(box[0] > 0) || !(box[1] > 0) || !(box[2] > 0)) {
    pgk_set_error("pgk_set_box: invalid box");
    return PGK_EINVAL;
  }
  for (int d = 0; d < 3; d++) c->box[d] = box[d];
  c->pme_grid_ready = false;
  return PGK_OK;
}

int pgk_set_positions(pgk_ctx *c, const double *pos) {
  if (!c || !pos) {
    pgk_set_error("pgk_set_positions: null argument");
    return PGK_EINVAL;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  HIP_CHECK(hipMemcpyAsync(c->d_pos, pos, sizeof(double) * 3 * c->natom, hipMemcpyHostToDevice,
                           c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->have_pos = true;
  c->hf_ready = false;   // calheatf pairs the arrays with the positions of their call
  c->pme_grid_ready = false;
  return PGK_OK;
}

int pgk_set_positions_device(pgk_ctx *c, const double *d_pos) {
  if (!c || !d_pos) {
    pgk_set_error("pgk_set_positions_device: null argument");
    return PGK_EINVAL;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  HIP_CHECK(hipMemcpyAsync(c->d_pos, d_pos, sizeof(double) * 3 * c->natom,
                           hipMemcpyDeviceToDevice, c->stream));
  c->have_pos = true;
  c->hf_ready = false;   // calheatf pairs the arrays with the positions of their call
  c->pme_grid_ready = false;
  return PGK_OK;
}

int pgk_build_list(pgk_ctx *c, double rcut_book) {
  if (!c || !(rcut_book > 0)) {
    pgk_set_error("pgk_build_list: invalid argument");
    return PGK_EINVAL;
  }
  for (int d = 0; d < 3; d++)
    if (rcut_book > 0.5 * c->box[d]) {
      // prepmd.F90:374-382 stops the same way
      pgk_set_error("pgk_build_list: rcut_book exceeds half the box");
      return PGK_EINVAL;
    }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  pgk_prof_begin(c, PGK_PROF_LIST);
  rc = pgk_list_build(c, rcut_book);
  pgk_prof_end(c, PGK_PROF_LIST);
  if (c->prof) {
    (void)hipStreamSynchronize(c->stream);
    pgk_prof_collect(c);
  }
  return rc;
}

int pgk_list_npairs(pgk_ctx *c, int64_t *np) {
  if (!c || !np) return PGK_EINVAL;
  if (!c->have_list) {
    pgk_set_error("pgk_list_npairs: no list");
    return PGK_ESTATE;
  }
  int rc = set_device_ctx(c);
  if (rc || (rc = pgk_list_count_pairs(c))) return rc;
  *np = c->npairs;
  return PGK_OK;
}

int pgk_list_pairs(pgk_ctx *c, int32_t *pairs, int64_t capacity) {
  if (!c || !pairs) return PGK_EINVAL;
  if (!c->have_list) {
    pgk_set_error("pgk_list_pairs: no list");
    return PGK_ESTATE;
  }
  int rc = set_device_ctx(c);
  if (rc || (rc = pgk_list_count_pairs(c))) return rc;
  if (capacity < c->npairs) {
    pgk_set_error("pgk_list_pairs: capacity too small");
    return PGK_EINVAL;
  }
  std::vector<int32_t> start(c->ncl + 1), cj(c->npair_cl), atoms((size_t)c->ncl * PGK_CL);
  std::vector<uint32_t> mask(c->npair_cl);
  HIP_CHECK(hipMemcpy(start.data(), c->d_ci_start, sizeof(int32_t) * (c->ncl + 1),
                      hipMemcpyDeviceToHost));
  if (c->npair_cl) {
    HIP_CHECK(hipMemcpy(cj.data(), c->d_cj, sizeof(int32_t) * c->npair_cl, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(mask.data(), c->d_mask, sizeof(uint32_t) * c->npair_cl,
                        hipMemcpyDeviceToHost));
  }
  HIP_CHECK(hipMemcpy(atoms.data(), c->d_cl_atom, sizeof(int32_t) * atoms.size(),
                      hipMemcpyDeviceToHost));
  std::vector<std::pair<int32_t, int32_t>> v;
  v.reserve(c->npairs);
  for (int ci = 0; ci < c->ncl; ci++)
    for (int e = start[ci]; e < start[ci + 1]; e++)
      for (int b = 0; b < 32; b++)
        if ((mask[e] >> b) & 1u) {
          // entry cj[e] = 2 cluster + half; bit b = ii + 8 jq
          int ai = atoms[(size_t)ci * PGK_CL + (b & 7)], aj = atoms[(size_t)cj[e] * 4 + (b >> 3)];
          v.emplace_back(std::min(ai, aj) + 1, std::max(ai, aj) + 1);
        }
  std::sort(v.begin(), v.end());
  for (size_t k = 0; k < v.size(); k++) {
    pairs[2 * k] = v[k].first;
    pairs[2 * k + 1] = v[k].second;
  }
  return PGK_OK;
}

int pgk_calforce(pgk_ctx *c, int32_t mts_flag, const pgk_mts *mts, int32_t ifcalpreatom,
                 pgk_force_out *out) {
  if (!c || !mts || !out || !out->force || !out->ene) {
    pgk_set_error("pgk_calforce: null argument");
    return PGK_EINVAL;
  }
  bool split = out->for_viri_coul && out->for_viri_lj;
  int rc = run_calforce(c, mts_flag, mts, ifcalpreatom, split, c->d_force, c->d_fvc, c->d_fvl,
                        out->ene, out->virt);
  if (rc) return rc;
  const size_t nb = sizeof(double) * 3 * c->natom;
  HIP_CHECK(hipMemcpy(out->force, c->d_force, nb, hipMemcpyDeviceToHost));
  if (split) {
    HIP_CHECK(hipMemcpy(out->for_viri_coul, c->d_fvc, nb, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(out->for_viri_lj, c->d_fvl, nb, hipMemcpyDeviceToHost));
  }
  return PGK_OK;
}

int pgk_calforce_device(pgk_ctx *c, int32_t mts_flag, const pgk_mts *mts, int32_t ifcalpreatom,
                        double *d_force, double *ene, double *virt) {
  if (!c || !mts || !ene) {
    pgk_set_error("pgk_calforce_device: null argument");
    return PGK_EINVAL;
  }
  return run_calforce(c, mts_flag, mts, ifcalpreatom, false, d_force ? d_force : c->d_force,
                      nullptr, nullptr, ene, virt);
}

// grow a device buffer to at least `need` bytes (contents not kept)
int grow_dev(void **p, size_t *cap, size_t need) {
  if (need <= *cap && *p) return PGK_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  HIP_CHECK(hipMalloc(p, need));
  *cap = need;
  return PGK_OK;
}

int pgk_calforce_async(pgk_ctx *c, int32_t mts_flag, const pgk_mts *mts, int32_t ifcalpreatom,
                       double *d_force) {
  if (!c || !mts) {
    pgk_set_error("pgk_calforce_async: null argument");
    return PGK_EINVAL;
  }
  return run_calforce(c, mts_flag, mts, ifcalpreatom, false, d_force ? d_force : c->d_force,
                      nullptr, nullptr, nullptr, nullptr, false, false);
}

int pgk_energies(pgk_ctx *c, double *ene, double *virt) {
  if (!c || !ene) {
    pgk_set_error("pgk_energies: null argument");
    return PGK_EINVAL;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  return finalize_energies(c, ene, virt);
}

// Shared set-up of the heat-flux calls: checks the description, uploads the
// regions, hftyp and molecule centres, clears the per-atom sums and points
// hfdev at them.  flux: no W arrays; the collision part is summed into
// d_hj_jcol / d_hj_jpart with the velocities already in d_hj_vel.
// dev_com: the molecule centres are computed on the device (k_calcom) rather
// than given by the caller
// checked: the description was validated before (the MD loop's samples of
// the description md_hf_begin checked), so the per-atom scans are skipped
static int hf_setup(pgk_ctx *c, const pgk_hf_desc *hf, bool flux, const char *who,
                    bool upload = true, bool dev_com = false, bool checked = false) {
  if (hf->mode < PGK_HF_VOLUME || hf->mode > PGK_HF_BULK || hf->nregion < 1 || !hf->z1 ||
      !hf->z2 || !hf->hftyp) {
    pgk_set_error(std::string(who) + ": invalid heat-flux description");
    return PGK_EINVAL;
  }
  const int n = c->natom, nreg = hf->nregion;
  bool need_com = false;
  int maxmol = 0;
  if (checked) {
    need_com = c->md_hf_mole;
    maxmol = c->hj_nmol;
  } else {
    for (int i = 0; i < n; i++) {
      if (hf->hftyp[i] != PGK_HFTYP_ATOM && hf->hftyp[i] != PGK_HFTYP_MOLE) {
        pgk_set_error(std::string(who) + ": hftyp must be ATOM (1) or MOLE (2)");
        return PGK_EINVAL;
      }
      if (hf->hftyp[i] == PGK_HFTYP_MOLE) need_com = true;
    }
    for (int i = 0; i < n; i++) maxmol = std::max(maxmol, c->molid[i]);
  }
  if (need_com && !dev_com && (!hf->molecom || hf->nmol < maxmol)) {
    pgk_set_error(std::string(who) + ": MOLE atoms need molecom for every molecule");
    return PGK_EINVAL;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  hipStream_t st = c->stream;
  const size_t nw = (hf->mode == PGK_HF_BULK) ? 9 : 3;
  const size_t pe_b = sizeof(double) * PGK_NHF * (size_t)n;
  const size_t w_b = sizeof(double) * PGK_NHF * (size_t)nreg * n * nw;
  const size_t wout_b = sizeof(double) * PGK_NHF * (size_t)nreg * n * 9;
  size_t typ_cap = c->d_hf_typ ? sizeof(int32_t) * n : 0;
  size_t pe_cap = c->d_hf_pe ? pe_b : 0, peo_cap = c->d_hf_pe_out ? pe_b : 0;
  if ((rc = grow_dev((void **)&c->d_hf_z, &c->hf_z_cap, sizeof(double) * 2 * nreg)) ||
      (rc = grow_dev((void **)&c->d_hf_typ, &typ_cap, sizeof(int32_t) * n)) ||
      (rc = grow_dev(&c->d_hf_pe, &pe_cap, pe_b)))
    return rc;
  if (!flux && ((rc = grow_dev(&c->d_hf_w, &c->hf_w_cap, w_b)) ||
                (rc = grow_dev((void **)&c->d_hf_pe_out, &peo_cap, pe_b)) ||
                (rc = grow_dev((void **)&c->d_hf_w_out, &c->hf_wout_cap, wout_b))))
    return rc;
  const int ncom = dev_com ? maxmol : hf->nmol;
  if (need_com &&
      (rc = grow_dev((void **)&c->d_hf_com, &c->hf_com_cap, sizeof(double) * 3 * ncom)))
    return rc;
  if (upload) {
    HIP_CHECK(hipMemcpyAsync(c->d_hf_z, hf->z1, sizeof(double) * nreg, hipMemcpyHostToDevice,
                             st));
    HIP_CHECK(hipMemcpyAsync(c->d_hf_z + nreg, hf->z2, sizeof(double) * nreg,
                             hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(c->d_hf_typ, hf->hftyp, sizeof(int32_t) * n,
                             hipMemcpyHostToDevice, st));
  }
  if (need_com && !dev_com)
    HIP_CHECK(hipMemcpyAsync(c->d_hf_com, hf->molecom, sizeof(double) * 3 * hf->nmol,
                             hipMemcpyHostToDevice, st));
  if (!flux) {
    HIP_CHECK(hipMemsetAsync(c->d_hf_pe, 0, pe_b, st));
    HIP_CHECK(hipMemsetAsync(c->d_hf_w, 0, w_b, st));
  }
  HfDev &h = c->hfdev;
  h.mode = hf->mode;
  h.nreg = nreg;
  h.natom = n;
  h.nw = (int)nw;
  h.z1 = c->d_hf_z;
  h.z2 = c->d_hf_z + nreg;
  {   // equal bins tiling [z1[0], z2[nreg-1]]
    const double w = hf->z2[0] - hf->z1[0];
    bool ub = nreg >= 2 && w > 0.0;
    for (int r = 0; ub && r < nreg; r++)
      ub = fabs(hf->z1[r] - (hf->z1[0] + r * w)) <= 1e-9 * w &&
           fabs(hf->z2[r] - hf->z1[r] - w) <= 1e-9 * w;
    h.ubin = ub ? 1 : 0;
    h.ubz0 = hf->z1[0];
    h.ubw = ub ? w : 1.0;
    h.ubwi = 1.0 / h.ubw;
    bool tiled = ub;
    for (int r = 0; tiled && r < nreg; r++)
      tiled = hf->z1[r] < hf->z2[r] && (r + 1 == nreg || hf->z2[r] == hf->z1[r + 1]);
    h.tiled = tiled ? 1 : 0;
  }
  h.hftyp = c->d_hf_typ;
  h.molid = c->d_molid;
  h.molecom = need_com ? c->d_hf_com : nullptr;
  h.pos = c->d_pos;
  for (int d = 0; d < 3; d++) {
    h.L[d] = c->box[d];
    h.Li[d] = 1.0 / c->box[d];
  }
  h.pe = c->d_hf_pe;
  h.w = flux ? nullptr : c->d_hf_w;
  h.flux = flux ? 1 : 0;
  h.vel = flux ? c->d_hj_vel : nullptr;
  h.tvel = nullptr;
  h.direct_ok = (flux && hf->mode == PGK_HF_BULK && nreg == 1 && !need_com &&
                 hf->z1[0] <= 0.0 && hf->z2[0] >= c->box[2])
                    ? 1
                    : 0;
  h.direct_tr = 0;
  h.jcol = nullptr;
  h.jpart = nullptr;
  if (flux) {
    const size_t jcol_b = sizeof(double) * PGK_NHF * (size_t)nreg * 3;
    const size_t nblk = (size_t)pgk_grid_blocks(std::max<int64_t>(pgk_pair_waves(c), 1), 4);
    // bulk: 12 flux sums and 12 transport sums per block; volume: [4][nreg]
    const size_t jpart_b =
        sizeof(double) * nblk * (hf->mode == PGK_HF_BULK ? 24 : 4 * (size_t)nreg);
    if ((rc = grow_dev((void **)&c->d_hj_jcol, &c->hj_jcol_cap, jcol_b)) ||
        (rc = grow_dev((void **)&c->d_hj_jpart, &c->hj_jpart_cap, jpart_b)))
      return rc;
    // the per-atom energies and both flux sums start at zero: one launch
    if ((rc = pgk_zero3_launch(st, c->d_hf_pe, pe_b / 8, c->d_hj_jcol, jcol_b / 8, c->d_hj_jpart,
                               jpart_b / 8)))
      return rc;
    h.jcol = c->d_hj_jcol;
    h.jpart = c->d_hj_jpart;
  }
  return PGK_OK;
}

static int copy_force_out(pgk_ctx *c, pgk_force_out *out) {
  const size_t nb = sizeof(double) * 3 * c->natom;
  HIP_CHECK(hipMemcpy(out->force, c->d_force, nb, hipMemcpyDeviceToHost));
  if (out->for_viri_coul && out->for_viri_lj) {
    HIP_CHECK(hipMemcpy(out->for_viri_coul, c->d_fvc, nb, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(out->for_viri_lj, c->d_fvl, nb, hipMemcpyDeviceToHost));
  }
  return PGK_OK;
}

int pgk_calforce_hf(pgk_ctx *c, int32_t mts_flag, const pgk_mts *mts, const pgk_hf_desc *hf,
                    pgk_force_out *out, pgk_hf_out *hfo) {
  if (!c || !mts || !hf || !out || !out->force || !out->ene) {
    pgk_set_error("pgk_calforce_hf: null argument");
    return PGK_EINVAL;
  }
  // per-atom arrays the caller does not take stay on the device only
  const bool want_pe = hfo && hfo->pot_atm, want_w = hfo && hfo->viri_atm;
  int rc = hf_setup(c, hf, false, "pgk_calforce_hf");
  if (rc) return rc;
  const size_t n = (size_t)c->natom, nreg = (size_t)hf->nregion;
  const size_t pe_b = sizeof(double) * PGK_NHF * n;
  const size_t wout_b = sizeof(double) * PGK_NHF * nreg * n * 9;
  const bool split = out->for_viri_coul && out->for_viri_lj;
  c->hf_ready = false;
  c->hf_export = want_pe || want_w;
  rc = run_calforce(c, mts_flag, mts, 1, split, c->d_force, c->d_fvc, c->d_fvl, out->ene,
                    out->virt, true);
  if (rc) return rc;
  c->hf_ready = true;
  if ((rc = copy_force_out(c, out))) return rc;
  if (want_pe) HIP_CHECK(hipMemcpy(hfo->pot_atm, c->d_hf_pe_out, pe_b, hipMemcpyDeviceToHost));
  if (want_w) HIP_CHECK(hipMemcpy(hfo->viri_atm, c->d_hf_w_out, wout_b, hipMemcpyDeviceToHost));
  return PGK_OK;
}

// host velocities / masses -> d_hj_vel / d_hj_mass (mass NULL keeps the last)
static int hj_upload(pgk_ctx *c, const double *vel, const double *mass, const char *who) {
  if (!mass && !c->hj_have_mass) {
    pgk_set_error(std::string(who) + ": no masses given yet");
    return PGK_EINVAL;
  }
  const int n = c->natom;
  int rc;
  if (!c->d_hj_vel) {
    std::vector<double> z((size_t)3 * n, 0.0);
    if ((rc = upload(&c->d_hj_vel, z))) return rc;
  }
  HIP_CHECK(hipMemcpyAsync(c->d_hj_vel, vel, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                           c->stream));
  if (mass) {
    if (!c->d_hj_mass) {
      std::vector<double> z(n, 0.0);
      if ((rc = upload(&c->d_hj_mass, z))) return rc;
    }
    HIP_CHECK(hipMemcpyAsync(c->d_hj_mass, mass, sizeof(double) * n, hipMemcpyHostToDevice,
                             c->stream));
    c->hj_have_mass = true;
  }
  return PGK_OK;
}

// the atoms of each molecule, in index order (for centre-of-mass velocities)
static int hj_molecules(pgk_ctx *c) {
  if (c->d_hj_mol_start) return PGK_OK;
  const int n = c->natom;
  int nmol = 0, rc;
  for (int i = 0; i < n; i++) nmol = std::max(nmol, c->molid[i]);
  std::vector<int32_t> start(nmol + 1, 0), atoms(std::max(n, 1));
  for (int i = 0; i < n; i++) start[c->molid[i]]++;
  for (int m = 0; m < nmol; m++) start[m + 1] += start[m];
  std::vector<int32_t> fill(start.begin(), start.end() - 1);
  for (int i = 0; i < n; i++) atoms[fill[c->molid[i] - 1]++] = i;
  if ((rc = upload(&c->d_hj_mol_start, start)) || (rc = upload(&c->d_hj_mol_atoms, atoms)))
    return rc;
  std::vector<double> mv((size_t)3 * std::max(nmol, 1), 0.0);
  if ((rc = upload(&c->d_hj_molevel, mv))) return rc;
  c->hj_nmol = nmol;
  return PGK_OK;
}

// the streaming velocities the standalone heat-flux calls subtract, or null
static const double *hj_vstrm(pgk_ctx *c) { return c->hf_vstrm_on ? c->d_hf_vstrm : nullptr; }

int pgk_set_atmvel_strm(pgk_ctx *c, const double *vstrm) {
  if (!c) {
    pgk_set_error("pgk_set_atmvel_strm: null context");
    return PGK_EINVAL;
  }
  if (!vstrm) {
    c->hf_vstrm_on = false;
    return PGK_OK;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  if ((rc = upload(&c->d_hf_vstrm,
                   std::vector<double>(vstrm, vstrm + (size_t)3 * c->natom))))
    return rc;
  c->hf_vstrm_on = true;
  return PGK_OK;
}

// calforce_hf and calheatf in one pass, without the per-atom W arrays
int pgk_calforce_heatf(pgk_ctx *c, int32_t mts_flag, const pgk_mts *mts, const pgk_hf_desc *hf,
                       const double *vel, const double *mass, pgk_force_out *out, double *hfj) {
  if (!c || !mts || !hf || !out || !out->force || !out->ene || !vel || !hfj) {
    pgk_set_error("pgk_calforce_heatf: null argument");
    return PGK_EINVAL;
  }
  if (hf->mode == PGK_HF_SURFACE) {
    pgk_set_error("pgk_calforce_heatf: the surface-based form is not offered");
    return PGK_EINVAL;
  }
  if (hf->mode == PGK_HF_VOLUME && hf->nregion > PGK_HEATF_MAXREG) {
    pgk_set_error("pgk_calforce_heatf: more than " + std::to_string(PGK_HEATF_MAXREG) +
                  " volume regions (use pgk_calforce_hf + pgk_calheatf)");
    return PGK_EINVAL;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  if ((rc = hj_upload(c, vel, mass, "pgk_calforce_heatf")) || (rc = hj_molecules(c)) ||
      (rc = hf_setup(c, hf, true, "pgk_calforce_heatf")) ||
      (rc = pgk_heatf_collision_vel(c, c->d_hj_vel, hj_vstrm(c), &c->hfdev.vel)))
    return rc;
  c->hfdev.tvel = c->d_hj_vel;   // what pgk_heatf_launch transports below
  c->hf_ready = false;
  c->hf_export = false;
  rc = run_calforce(c, mts_flag, mts, 1, out->for_viri_coul && out->for_viri_lj, c->d_force,
                    c->d_fvc, c->d_fvl, out->ene, out->virt, true);
  if (rc) return rc;
  if ((rc = pgk_heatf_launch(c, c->d_hj_vel, c->d_hj_mass, hj_vstrm(c)))) return rc;
  if ((rc = copy_force_out(c, out))) return rc;
  HIP_CHECK(hipMemcpyAsync(hfj, c->d_hj_out, sizeof(double) * (size_t)hf->nregion * PGK_NHFJ * 3,
                           hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->hfdev.flux = 0;
  return PGK_OK;
}

// calheatf over the arrays of the last pgk_calforce_hf call
int pgk_calheatf(pgk_ctx *c, const double *vel, const double *mass, double *hf) {
  if (!c || !vel || !hf) {
    pgk_set_error("pgk_calheatf: null argument");
    return PGK_EINVAL;
  }
  if (!c->hf_ready) {
    pgk_set_error("pgk_calheatf: no pgk_calforce_hf results on the device");
    return PGK_ESTATE;
  }
  if (c->hfdev.mode == PGK_HF_SURFACE) {
    pgk_set_error("pgk_calheatf: the surface-based form is not offered");
    return PGK_EINVAL;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  if ((rc = hj_upload(c, vel, mass, "pgk_calheatf")) || (rc = hj_molecules(c))) return rc;
  if ((rc = pgk_heatf_launch(c, c->d_hj_vel, c->d_hj_mass, hj_vstrm(c)))) return rc;
  HIP_CHECK(hipMemcpyAsync(hf, c->d_hj_out,
                           sizeof(double) * (size_t)c->hfdev.nreg * PGK_NHFJ * 3,
                           hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return PGK_OK;
}

// ---- device-resident NVE step ---------------------------------------------
int pgk_md_setup(pgk_ctx *c, const pgk_md_desc *md) {
  if (!c || !md || !md->mass || !(md->dt > 0) || md->ncons < 0 || md->ngroup < 0 ||
      (md->ncons > 0 && (!md->cons_ij || !md->cons_d || !md->group_start))) {
    pgk_set_error("pgk_md_setup: invalid description");
    return PGK_EINVAL;
  }
  const int n = c->natom;
  std::vector<double> minv(n);
  for (int i = 0; i < n; i++) {
    if (!(md->mass[i] > 0)) {
      pgk_set_error("pgk_md_setup: masses must be positive");
      return PGK_EINVAL;
    }
    minv[i] = 1.0 / md->mass[i];
  }
  std::vector<int32_t> gs, cij;
  std::vector<double> d2;
  if (md->ncons > 0) {
    gs.assign(md->group_start, md->group_start + md->ngroup + 1);
    if (gs[0] != 0 || gs[md->ngroup] != md->ncons) {
      pgk_set_error("pgk_md_setup: group_start must run from 0 to ncons");
      return PGK_EINVAL;
    }
    for (int g = 0; g < md->ngroup; g++)
      if (gs[g + 1] < gs[g]) {
        pgk_set_error("pgk_md_setup: group_start not monotone");
        return PGK_EINVAL;
      }
    for (int k = 0; k < md->ncons; k++) {
      const int i = md->cons_ij[2 * k], j = md->cons_ij[2 * k + 1];
      if (i < 1 || i > n || j < 1 || j > n || i == j || !(md->cons_d[k] > 0)) {
        pgk_set_error("pgk_md_setup: constraint out of range");
        return PGK_EINVAL;
      }
      cij.push_back(i - 1);
      cij.push_back(j - 1);
      d2.push_back(md->cons_d[k] * md->cons_d[k]);
    }
    // groups are independent molecules: one thread updates each group, so no
    // atom may belong to two of them
    std::vector<int32_t> owner(n, -1);
    for (int g = 0; g < md->ngroup; g++)
      for (int k = gs[g]; k < gs[g + 1]; k++)
        for (int e = 0; e < 2; e++) {
          int &o = owner[cij[2 * k + e]];
          if (o >= 0 && o != g) {
            pgk_set_error("pgk_md_setup: an atom belongs to two constraint groups");
            return PGK_EINVAL;
          }
          o = g;
        }
  }
  // triangle groups (3 atoms, 3 constraints) keep the host's constraint order;
  // their atoms are labelled a = the atom shared by the first two constraints,
  // b and c the others, so the order reads (a,b), (a,c), (b,c).  Swapping the
  // two atoms of a constraint changes no result (both vectors change sign).
  const int ng = md->ncons > 0 ? md->ngroup : 0;
  std::vector<int32_t> small(std::max(ng, 1), 0);
  for (int g = 0; g < ng; g++) {
    if (gs[g + 1] - gs[g] != 3) continue;
    const int k0 = gs[g];
    const int x0 = cij[2 * k0], y0 = cij[2 * k0 + 1];
    const int x1 = cij[2 * k0 + 2], y1 = cij[2 * k0 + 3];
    const int x2 = cij[2 * k0 + 4], y2 = cij[2 * k0 + 5];
    const int a = (x0 == x1 || x0 == y1) ? x0 : ((y0 == x1 || y0 == y1) ? y0 : -1);
    if (a < 0) continue;
    const int b = (x0 == a) ? y0 : x0, cc = (x1 == a) ? y1 : x1;
    if (b == a || cc == a || cc == b) continue;
    if (!((x2 == b && y2 == cc) || (x2 == cc && y2 == b))) continue;
    const int lab[3][2] = {{a, b}, {a, cc}, {b, cc}};
    for (int q = 0; q < 3; q++) {
      cij[2 * (k0 + q)] = lab[q][0];
      cij[2 * (k0 + q) + 1] = lab[q][1];
    }
    small[g] = 1;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  if ((rc = upload(&c->d_md_minv, minv)) || (rc = upload(&c->d_md_gstart, gs)) ||
      (rc = upload(&c->d_md_cij, cij)) || (rc = upload(&c->d_md_d2, d2)) ||
      (rc = upload(&c->d_md_small, small)))
    return rc;
  std::vector<double> zero((size_t)3 * n, 0.0);
  std::vector<int32_t> one(1, 0);
  if ((rc = upload(&c->d_md_vel, zero)) || (rc = upload(&c->d_md_pos_old, zero)) ||
      (rc = upload(&c->d_md_fail, one)))
    return rc;
  std::vector<double> rec((size_t)pgk_grid_blocks(n, 256) * PGK_REC, 0.0);
  if ((rc = upload(&c->d_md_ke, rec)) || (rc = upload(&c->d_md_kenhc, rec))) return rc;
  // every group a triangle: the half steps run as one kernel each, with the
  // atoms outside every group listed for their own threads
  bool all_tri = true;
  for (int g = 0; g < ng; g++) all_tri = all_tri && small[g];
  std::vector<int32_t> in_group(n, 0), free_atoms;
  for (size_t k = 0; k < cij.size(); k++) in_group[cij[k]] = 1;
  for (int i = 0; i < n; i++)
    if (!in_group[i]) free_atoms.push_back(i);
  if (free_atoms.empty()) free_atoms.push_back(0);   // upload needs one element
  if ((rc = upload(&c->d_md_free, free_atoms))) return rc;
  c->md_nfree = 0;
  for (int i = 0; i < n; i++) c->md_nfree += in_group[i] ? 0 : 1;
  c->md_all_tri = all_tri;
  c->md_ngroup = md->ncons > 0 ? md->ngroup : 0;
  c->md_ncons = md->ncons;
  c->md_dt = md->dt;
  c->md_tol = md->tol > 0 ? md->tol : 1e-10;
  c->md_mass.assign(md->mass, md->mass + n);
  c->md_ready = true;
  return PGK_OK;
}

int pgk_md_set_velocities(pgk_ctx *c, const double *vel) {
  if (!c || !vel || !c->md_ready) {
    pgk_set_error("pgk_md_set_velocities: no MD set-up or null argument");
    return PGK_EINVAL;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  HIP_CHECK(hipMemcpyAsync(c->d_md_vel, vel, sizeof(double) * 3 * c->natom,
                           hipMemcpyHostToDevice, c->stream));
  if ((rc = pgk_md_rattle_v(c))) return rc;
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return PGK_OK;
}

int pgk_md_get_state(pgk_ctx *c, double *pos, double *vel) {
  if (!c || !c->md_ready) {
    pgk_set_error("pgk_md_get_state: no MD set-up");
    return PGK_EINVAL;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  const size_t nb = sizeof(double) * 3 * c->natom;
  if (pos) HIP_CHECK(hipMemcpyAsync(pos, c->d_pos, nb, hipMemcpyDeviceToHost, c->stream));
  if (vel) HIP_CHECK(hipMemcpyAsync(vel, c->d_md_vel, nb, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return PGK_OK;
}

static int md_sample(pgk_ctx *c, double *row) {
  int rc = pgk_md_ke(c, c->d_md_ke);
  if (rc) return rc;
  // RESPA: the long call's record over the short call's, as calforce leaves
  // the terms it does not evaluate alone and sums every term into pot_tot
  double ene[PGK_NENE] = {};
  if (c->md_nshort > 1)
    for (int k = 0; k < PGK_NENE; k++) ene[k] = c->md_e_short[k];
  if ((rc = finalize_energies(c, ene, nullptr))) return rc;   // waits for the stream
  const int nb = pgk_grid_blocks(c->natom, 256);
  std::vector<double> rec((size_t)nb * PGK_REC);
  HIP_CHECK(hipMemcpy(rec.data(), c->d_md_ke, sizeof(double) * rec.size(),
                      hipMemcpyDeviceToHost));
  double ke = 0.0;
  for (int b = 0; b < nb; b++) ke += rec[(size_t)b * PGK_REC];
  int32_t fail = 0;
  HIP_CHECK(hipMemcpy(&fail, c->d_md_fail, sizeof(fail), hipMemcpyDeviceToHost));
  if (fail) {
    pgk_set_error("pgk_md_run: constraints did not converge (RATTLE)");
    return PGK_EOVERFLOW;
  }
  // with the chain on, E is the conserved quantity (calkin.F90:189-201)
  double eth = 0.0;
  const int m = c->md_nhc_m;
  if (m > 0) {
    std::vector<double> st((size_t)3 * m);
    HIP_CHECK(hipMemcpy(st.data(), c->d_md_nhc, sizeof(double) * st.size(),
                        hipMemcpyDeviceToHost));
    for (int i = 0; i < m; i++) eth += 0.5 * st[2 * m + i] * st[m + i] * st[m + i];
    eth += c->md_nhc_gnkt * st[0];
    for (int i = 1; i < m; i++) eth += c->md_nhc_gkt * st[i];
  }
  const double pe = ene[PGK_E_TOT];
  row[0] = ke;
  row[1] = pe;
  row[2] = ke + pe + eth;
  return PGK_OK;
}

static int hj_molecules(pgk_ctx *c);

// ---- heat-flux time series of the MD loop ----------------------------------
// the bonded terms (their heat-flux parts) and the pair heat-flux pass; the
// pair and reciprocal forces of the step are already known
const pgk_mts kMdHfMts = {PGK_MTS_LONG, PGK_MTS_LONG, PGK_MTS_LONG, PGK_MTS_LONG,
                          0, 0, PGK_MTS_LONG};

// size the series for nstep steps and upload the regions once
static int md_hf_begin(pgk_ctx *c, int nstep) {
  c->md_hf_rows = 0;
  if (c->md_hf_interval <= 0) return PGK_OK;
  const int nreg = (int)c->md_hf_z1.size();
  const size_t rows = (size_t)nstep / c->md_hf_interval + 1;
  int rc;
  if ((rc = grow_dev((void **)&c->d_md_hfseries, &c->md_hfseries_cap,
                     sizeof(double) * rows * nreg * PGK_NHFJ * 3)))
    return rc;
  if (!c->d_hj_fscratch) {
    std::vector<double> z((size_t)3 * c->natom, 0.0);
    if ((rc = upload(&c->d_hj_fscratch, z))) return rc;
  }
  if (c->md_hf_mole && (rc = hj_molecules(c))) return rc;
  const pgk_hf_desc d = {c->md_hf_mode, nreg, c->md_hf_z1.data(), c->md_hf_z2.data(),
                         c->md_hf_typ.data(), 0, nullptr};
  return hf_setup(c, &d, true, "pgk_md_run", true, true);
}

// at step s (a multiple of the interval): the fused heat-flux pass at the
// step's end state (positions x(t+dt), velocities v(t+dt)); its forces go to
// a scratch buffer, the row to the series
static int md_hf_sample(pgk_ctx *c, int s) {
  if (c->md_hf_interval <= 0 || s % c->md_hf_interval != 0) return PGK_OK;
  const int nreg = (int)c->md_hf_z1.size();
  const pgk_hf_desc d = {c->md_hf_mode, nreg, c->md_hf_z1.data(), c->md_hf_z2.data(),
                         c->md_hf_typ.data(), 0, nullptr};
  int rc;
  if ((rc = hf_setup(c, &d, true, "pgk_md_run", false, true, true))) return rc;
  // MOLE centres: the molecules' centres of mass at this state (calcom.F90)
  if (c->md_hf_mole && (rc = pgk_calcom_launch(c, c->d_md_hfmass, c->d_hf_com))) return rc;
  // with streaming slabs the collision part and the kinetic energies are
  // taken relative to the streaming velocities (calheatf.F90:368-369)
  const double *vs = c->md_lt_nstrm ? c->d_md_vstrm : nullptr;
  if ((rc = pgk_heatf_collision_vel(c, c->d_md_vel, vs, &c->hfdev.vel))) return rc;
  c->hfdev.tvel = c->d_md_vel;   // what pgk_heatf_launch transports below
  c->hf_export = false;
  c->hf_pairs_only = true;
  // the energy records of the step stay the ones pgk_energies reports
  const int last_nseg = c->last_nseg, last_flag = c->last_mts_flag;
  int last_kind[16];
  memcpy(last_kind, c->last_seg_kind, sizeof(last_kind));
  const pgk_mts last_mts = c->last_mts;
  const bool last_pending = c->last_pending;
  rc = run_calforce(c, PGK_MTS_LONG, &kMdHfMts, 0, false, c->d_hj_fscratch, nullptr, nullptr,
                    nullptr, nullptr, true, false, false);
  c->hf_pairs_only = false;
  c->last_nseg = last_nseg;
  c->last_mts_flag = last_flag;
  memcpy(c->last_seg_kind, last_kind, sizeof(last_kind));
  c->last_mts = last_mts;
  c->last_pending = last_pending;
  if (rc) return rc;
  // the rows go straight to the series
  const size_t row = (size_t)nreg * PGK_NHFJ * 3;
  if ((rc = pgk_heatf_launch(c, c->d_md_vel, c->d_md_hfmass, vs,
                             c->d_md_hfseries + (size_t)c->md_hf_rows * row)))
    return rc;
  c->hfdev.flux = 0;
  c->md_hf_rows++;
  return PGK_OK;
}

namespace {

const pgk_mts kMdMts = {PGK_MTS_LONG, PGK_MTS_LONG, PGK_MTS_LONG, PGK_MTS_LONG,
                        PGK_MTS_LONG, PGK_MTS_LONG, PGK_MTS_LONG};

// One context's MD loop, step by step.  A plain step (no list rebuild, no
// sample, no stage timing) launches the same kernels with the same arguments
// every time; with graphs on it is captured once into a HIP graph, after one
// plain step has run eagerly since the last rebuild (so no set-up work is
// left to do inside the capture), and then replayed.  A rebuild can
// reallocate the list buffers, so it drops the graph.
struct MdRunner {
  pgk_ctx *c;
  bool graphs;
  hipGraphExec_t gexec = nullptr;
  bool primed = false;
  MdRunner(pgk_ctx *ctx, bool use_graphs) : c(ctx), graphs(use_graphs) {}
  ~MdRunner() { drop(); }
  void drop() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    primed = false;
  }
  // With the Nose-Hoover chain on, a step opens and closes with nhcint.
  // Both calls see the same atom velocities (the chain moves only its own
  // variables; the atoms feel it through the half kicks), so the closing
  // call's kinetic-energy records serve the next step's opening call, and
  // begin() writes them for the first.
  // With the region Langevin thermostat as well, the chain couples to the
  // atoms outside every region; the region half that opens the step has just
  // changed the others, and a registration may have moved atoms between the
  // sets, so the opening call takes fresh records of the chain's atoms.
  int nhc_open() {
    int r;
    if (!c->md_nhc_m) return PGK_OK;
    if (c->md_lt_nreg && (r = pgk_md_ke(c, c->d_md_kenhc, true))) return r;
    return pgk_md_nhc_half(c);
  }
  int nhc_close() {
    int r;
    if (!c->md_nhc_m) return PGK_OK;
    if ((r = pgk_md_ke(c, c->d_md_kenhc, true))) return r;
    return pgk_md_nhc_half(c);
  }
  int eager(bool ene) {
    int r;
    if ((r = nhc_open()) || (r = pgk_md_kick_drift(c, c->d_force, c->md_dt, nullptr, 0.0)))
      return r;
    if ((r = run_calforce(c, PGK_MTS_LONG, &kMdMts, 0, false, c->d_force, nullptr, nullptr,
                          nullptr, nullptr, false, false, ene)))
      return r;
    if ((r = pgk_md_kick(c, c->d_force, c->md_dt, nullptr, 0.0))) return r;
    return nhc_close();
  }
  // Region Langevin thermostat: its first half opens the step, before the
  // first half kick; at the close the atoms are registered at the new
  // positions and the second half follows RATTLE (moldyn_hf.F90:1667-1678,
  // 2348-2372).  Steps with it are never replayed from a graph: every step
  // draws new random numbers.
  int lt_open() { return c->md_lt_nreg ? pgk_md_lt_half(c, true) : PGK_OK; }
  int lt_close() {
    int r;
    if (!c->md_lt_nreg) return PGK_OK;
    if ((r = pgk_md_lt_register(c))) return r;
    return pgk_md_lt_half(c, false);
  }
  bool respa() const { return c->md_nshort > 1; }
  // the short-level forces into d_md_fshort; at a sample their energy record
  // is finalised into md_e_short (the long call's record is the step's own)
  int short_forces(bool ene) {
    if (!ene)
      return run_calforce(c, PGK_MTS_SHORT, &c->md_mts, 0, false, c->d_md_fshort, nullptr,
                          nullptr, nullptr, nullptr, false, false, false);
    for (double &e : c->md_e_short) e = 0.0;
    return run_calforce(c, PGK_MTS_SHORT, &c->md_mts, 0, false, c->d_md_fshort, nullptr,
                        nullptr, c->md_e_short, nullptr, false, true, true);
  }
  // One RESPA step (moldyn.F90:1465-1771): nshort inner steps of dt/nshort;
  // the long half kick opens the first and closes the last, the long forces
  // are evaluated once, after the last drift.  A list rebuild happens at the
  // positions of the first short force call.
  int respa_step(bool rebuild, bool ene, double rcut_book) {
    int r;
    const int ns = c->md_nshort;
    const double dts = c->md_dt / ns;
    if ((r = lt_open())) return r;
    for (int s = 1; s <= ns; s++) {
      const bool last = s == ns;
      if ((r = pgk_md_kick_drift(c, c->d_md_fshort, dts, s == 1 ? c->d_force : nullptr,
                                 c->md_dt)))
        return r;
      if (s == 1 && rebuild) {
        drop();
        if ((r = pgk_build_list(c, rcut_book))) return r;
      }
      if ((r = short_forces(ene && last))) return r;
      if (last && (r = run_calforce(c, PGK_MTS_LONG, &c->md_mts, 0, false, c->d_force, nullptr,
                                    nullptr, nullptr, nullptr, false, false, ene)))
        return r;
      if ((r = pgk_md_kick(c, c->d_md_fshort, dts, last ? c->d_force : nullptr, c->md_dt)))
        return r;
    }
    return lt_close();
  }
  // step 0: list and forces at the starting positions
  int begin(double rcut_book) {
    int r;
    if ((r = pgk_build_list(c, rcut_book))) return r;
    if (c->md_lt_nreg) {
      HIP_CHECK(hipMemsetAsync(
          c->d_lt_det, 0,
          sizeof(double) * PGK_LT_MAXREG * (size_t)pgk_grid_blocks(c->natom, 256), c->stream));
      if ((r = pgk_md_lt_register(c))) return r;
    }
    if ((r = pgk_md_strm_assign(c))) return r;
    if (c->md_nhc_m && (r = pgk_md_ke(c, c->d_md_kenhc, true))) return r;
    if (respa()) {
      if ((r = short_forces(true))) return r;
      return run_calforce(c, PGK_MTS_LONG, &c->md_mts, 0, false, c->d_force, nullptr, nullptr,
                          nullptr, nullptr, false, false);
    }
    return run_calforce(c, PGK_MTS_LONG, &kMdMts, 0, false, c->d_force, nullptr, nullptr,
                        nullptr, nullptr, false, false);
  }
  int step(bool rebuild, bool ene, double rcut_book) {
    int r;
    if (respa()) return respa_step(rebuild, ene, rcut_book);
    const bool plain = !rebuild && !ene && !c->prof && !c->acc_dirty && graphs && !c->md_lt_nreg;
    if (plain && gexec) {
      HIP_CHECK(hipGraphLaunch(gexec, c->stream));
      return PGK_OK;
    }
    if (plain && primed) {
      hipGraph_t graph = nullptr;
      HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int rcap = eager(false);
      const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
      if (rcap || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (!rcap)
          pgk_set_error(std::string("MD step: graph capture failed: ") + hipGetErrorString(ec));
        return rcap ? rcap : PGK_EDEVICE;
      }
      const hipError_t ei = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        gexec = nullptr;
        pgk_set_error(std::string("MD step: graph instantiation failed: ") +
                      hipGetErrorString(ei));
        return PGK_EDEVICE;
      }
      HIP_CHECK(hipGraphLaunch(gexec, c->stream));
      return PGK_OK;
    }
    if ((r = lt_open()) || (r = nhc_open()) ||
        (r = pgk_md_kick_drift(c, c->d_force, c->md_dt, nullptr, 0.0)))
      return r;
    if (rebuild) {
      drop();
      if ((r = pgk_build_list(c, rcut_book))) return r;
    }
    if ((r = run_calforce(c, PGK_MTS_LONG, &kMdMts, 0, false, c->d_force, nullptr, nullptr,
                          nullptr, nullptr, false, false, ene)))
      return r;
    if ((r = pgk_md_kick(c, c->d_force, c->md_dt, nullptr, 0.0)) || (r = nhc_close()) ||
        (r = lt_close()))
      return r;
    if (plain) primed = true;
    return PGK_OK;
  }
};

}  // namespace

int pgk_md_run(pgk_ctx *c, int32_t nstep, int32_t nstep_book, double rcut_book,
               int32_t nsample, double *samples) {
  if (!c || !c->md_ready || nstep < 0 || nstep_book < 1 || nsample < 1 || !samples) {
    pgk_set_error("pgk_md_run: no MD set-up or invalid argument");
    return PGK_EINVAL;
  }
  if (!c->have_pos) {
    pgk_set_error("pgk_md_run: positions not set");
    return PGK_ESTATE;
  }
  if (c->md_nshort > 1 && c->md_nhc_m) {
    pgk_set_error("pgk_md_run: RESPA with the Nose-Hoover chain is not offered");
    return PGK_EINVAL;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  c->hf_ready = false;
  MdRunner run(c, c->md_graphs);
  if ((rc = md_hf_begin(c, nstep)) || (rc = run.begin(rcut_book)) ||
      (rc = md_sample(c, samples)) || (rc = md_hf_sample(c, 0)))
    return rc;
  int row = 1;
  for (int s = 1; s <= nstep; s++) {
    // energies only where a sample (or the caller, after the last step) reads them
    const bool ene = (s % nsample == 0) || s == nstep;
    if ((rc = run.step(s % nstep_book == 0, ene, rcut_book))) return rc;
    if (s % nsample == 0) {
      if ((rc = md_sample(c, samples + 3 * row))) return rc;
      row++;
    }
    if ((rc = md_hf_sample(c, s))) return rc;
  }
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return PGK_OK;
}

// Many independent replicas, each in its own context with its own stream:
// a host thread queues step s of each of its replicas before step s+1 of
// any, so the GPU runs the small replicas' kernels side by side instead of
// one system at a time
// (plain steps can be replayed from one captured graph per replica, see
// pgk_md_use_graphs).  Each replica runs exactly the single-context step, so
// its trajectory is the one pgk_md_run gives.
int pgk_md_run_batch(pgk_ctx *const *ctxs, int32_t n, int32_t nstep, int32_t nstep_book,
                     const double *rcut_book, int32_t nsample, double *samples) {
  if (!ctxs || n < 1 || nstep < 0 || nstep_book < 1 || nsample < 1 || !samples || !rcut_book) {
    pgk_set_error("pgk_md_run_batch: invalid argument");
    return PGK_EINVAL;
  }
  for (int r = 0; r < n; r++) {
    pgk_ctx *c = ctxs[r];
    if (c && c->md_nshort > 1 && c->md_nhc_m) {
      pgk_set_error("pgk_md_run_batch: RESPA with the Nose-Hoover chain is not offered");
      return PGK_EINVAL;
    }
    if (!c || !c->md_ready || !c->have_pos) {
      pgk_set_error("pgk_md_run_batch: replica " + std::to_string(r) +
                    " has no MD set-up or no positions");
      return PGK_ESTATE;
    }
    for (int q = 0; q < r; q++)
      if (ctxs[q] == c) {
        pgk_set_error("pgk_md_run_batch: a context appears twice");
        return PGK_EINVAL;
      }
  }
  const int nrow = nstep / nsample + 1;
  // Host threads share the replicas (largest first, greedy by atom count),
  // each queuing its replicas' steps in turn, so launch overhead and the
  // host waits of list rebuilds and samples overlap across threads.
  // (With graph replay on, one thread: a capture on one replica's stream
  // must not overlap another thread's synchronous copies.)
  bool any_graphs = false;
  for (int r = 0; r < n; r++) {
    any_graphs = any_graphs || ctxs[r]->md_batch_graphs;
    ctxs[r]->hf_ready = false;
  }
  const int nth = any_graphs ? 1 : std::min(n, 8);
  std::vector<std::vector<int>> part(nth);
  {
    std::vector<int> order(n);
    for (int r = 0; r < n; r++) order[r] = r;
    std::sort(order.begin(), order.end(),
              [&](int x, int y) { return ctxs[x]->natom > ctxs[y]->natom; });
    std::vector<int64_t> load(nth, 0);
    for (int r : order) {
      const int t = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      part[t].push_back(r);
      load[t] += ctxs[r]->natom;
    }
  }
  std::vector<int> trc(nth, PGK_OK);
  std::vector<std::string> terr(nth);
  auto worker = [&](int t) {
    const std::vector<int> &mine = part[t];
    std::vector<std::unique_ptr<MdRunner>> run;
    for (int r : mine) run.emplace_back(new MdRunner(ctxs[r], ctxs[r]->md_batch_graphs));
    auto fail = [&](int rc) {
      trc[t] = rc;
      terr[t] = pgk_last_error();
    };
    int rc;
    for (size_t k = 0; k < mine.size(); k++) {
      const int r = mine[k];
      if ((rc = set_device_ctx(ctxs[r])) || (rc = md_hf_begin(ctxs[r], nstep)) ||
          (rc = run[k]->begin(rcut_book[r])) ||
          (rc = md_sample(ctxs[r], samples + (size_t)3 * nrow * r)) ||
          (rc = md_hf_sample(ctxs[r], 0)))
        return fail(rc);
    }
    int row = 1;
    for (int s = 1; s <= nstep; s++) {
      const bool rebuild = (s % nstep_book == 0);
      const bool ene = (s % nsample == 0) || s == nstep;
      for (size_t k = 0; k < mine.size(); k++) {
        const int r = mine[k];
        if ((rc = set_device_ctx(ctxs[r])) || (rc = run[k]->step(rebuild, ene, rcut_book[r])))
          return fail(rc);
      }
      if (s % nsample == 0) {
        for (int r : mine)
          if ((rc = md_sample(ctxs[r], samples + (size_t)3 * nrow * r + 3 * row)))
            return fail(rc);
        row++;
      }
      for (int r : mine)
        if ((rc = set_device_ctx(ctxs[r])) || (rc = md_hf_sample(ctxs[r], s))) return fail(rc);
    }
    for (int r : mine) {
      const hipError_t e = hipStreamSynchronize(ctxs[r]->stream);
      if (e != hipSuccess) {
        pgk_set_error(std::string("pgk_md_run_batch: ") + hipGetErrorString(e));
        return fail(PGK_EDEVICE);
      }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nth; t++) th.emplace_back(worker, t);
  worker(0);
  for (auto &x : th) x.join();
  for (int t = 0; t < nth; t++)
    if (trc[t] != PGK_OK) {
      pgk_set_error(terr[t]);
      return trc[t];
    }
  return PGK_OK;
}

int pgk_md_nhc(pgk_ctx *c, const pgk_nhc_desc *nhc) {
  if (!c || !c->md_ready) {
    pgk_set_error("pgk_md_nhc: no MD set-up");
    return PGK_EINVAL;
  }
  if (!nhc) {
    c->md_nhc_m = 0;
    return PGK_OK;
  }
  const int m = nhc->mchain;
  if (m < 2 || m > 10 || !(nhc->temperature > 0) || !(nhc->tfreq > 0) || !(nhc->nfree > 0) ||
      (nhc->nyosh != 1 && nhc->nyosh != 3 && nhc->nyosh != 5) || nhc->next < 1) {
    pgk_set_error("pgk_md_nhc: need 2 <= mchain <= 10, temperature, tfreq, nfree > 0, "
                  "nyosh in {1, 3, 5}, next >= 1");
    return PGK_EINVAL;
  }
  // kT in eref, masses in eref fs^2 (prepmd.F90:515-518)
  const double gkt = 1.380649e-23 / 6.9511042e-21 * nhc->temperature;
  std::vector<double> st((size_t)3 * m + 2, 0.0);
  for (int i = 0; i < m; i++) {
    st[i] = nhc->xlogs ? nhc->xlogs[i] : 0.0;
    st[m + i] = nhc->vlogs ? nhc->vlogs[i] : 0.0;
    st[2 * m + i] = (i == 0 ? nhc->nfree * gkt : gkt) / nhc->tfreq / nhc->tfreq;
  }
  // Yoshida-Suzuki weights (prepmd.F90:534-548)
  double ys[5] = {1.0, 0.0, 0.0, 0.0, 0.0};
  if (nhc->nyosh == 3) {
    ys[0] = ys[2] = 1.0 / (2.0 - pow(2.0, 1.0 / 3.0));
    ys[1] = 1.0 - 2.0 * ys[0];
  } else if (nhc->nyosh == 5) {
    ys[0] = ys[1] = ys[3] = ys[4] = 1.0 / (4.0 - pow(4.0, 1.0 / 3.0));
    ys[2] = 1.0 - 4.0 * ys[0];
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  if ((rc = upload(&c->d_md_nhc, st))) return rc;
  c->md_nhc_m = m;
  c->md_nhc_nyosh = nhc->nyosh;
  c->md_nhc_next = nhc->next;
  c->md_nhc_gkt = gkt;
  c->md_nhc_gnkt = gkt * nhc->nfree;
  for (int k = 0; k < 5; k++) c->md_nhc_ys[k] = ys[k];
  return PGK_OK;
}

int pgk_md_nhc_state(pgk_ctx *c, double *xlogs, double *vlogs) {
  if (!c || !c->md_nhc_m) {
    pgk_set_error("pgk_md_nhc_state: no thermostat set up");
    return PGK_EINVAL;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  HIP_CHECK(hipStreamSynchronize(c->stream));
  const int m = c->md_nhc_m;
  std::vector<double> st((size_t)2 * m);
  HIP_CHECK(hipMemcpy(st.data(), c->d_md_nhc, sizeof(double) * st.size(),
                      hipMemcpyDeviceToHost));
  for (int i = 0; i < m; i++) {
    if (xlogs) xlogs[i] = st[i];
    if (vlogs) vlogs[i] = st[m + i];
  }
  return PGK_OK;
}

int pgk_recip_energy_atom(pgk_ctx *c, double *e_atom) {
  if (!c || !e_atom) {
    pgk_set_error("pgk_recip_energy_atom: null argument");
    return PGK_EINVAL;
  }
  if (!c->pme_grid_ready) {
    pgk_set_error("pgk_recip_energy_atom: no reciprocal-space evaluation at the current "
                  "positions");
    return PGK_ESTATE;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  if (!c->d_hj_fscratch) {   // natom doubles are enough; the buffer holds 3 natom
    std::vector<double> z((size_t)3 * c->natom, 0.0);
    if ((rc = upload(&c->d_hj_fscratch, z))) return rc;
  }
  if ((rc = pgk_pme_atom_energy(c, c->d_hj_fscratch))) return rc;
  HIP_CHECK(hipMemcpyAsync(e_atom, c->d_hj_fscratch, sizeof(double) * c->natom,
                           hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return PGK_OK;
}

int pgk_recip_virial_atom(pgk_ctx *c, double *viri_atm) {
  if (!c || !viri_atm) {
    pgk_set_error("pgk_recip_virial_atom: null argument");
    return PGK_EINVAL;
  }
  if (!c->pme_grid_ready) {
    pgk_set_error("pgk_recip_virial_atom: no reciprocal-space evaluation at the current "
                  "positions");
    return PGK_ESTATE;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  std::vector<double> z((size_t)9 * c->natom, 0.0);
  double *d_w = nullptr;
  if ((rc = upload(&d_w, z))) return rc;
  if ((rc = pgk_pme_atom_virial(c, d_w))) {
    (void)hipFree(d_w);
    return rc;
  }
  const hipError_t e1 = hipMemcpyAsync(viri_atm, d_w, sizeof(double) * z.size(),
                                       hipMemcpyDeviceToHost, c->stream);
  const hipError_t e2 = hipStreamSynchronize(c->stream);
  (void)hipFree(d_w);
  HIP_CHECK(e1);
  HIP_CHECK(e2);
  return PGK_OK;
}

int pgk_md_respa(pgk_ctx *c, int32_t nstep_short, const pgk_mts *mts) {
  if (!c || !c->md_ready || nstep_short < 1) {
    pgk_set_error("pgk_md_respa: no MD set-up or nstep_short < 1");
    return PGK_EINVAL;
  }
  pgk_mts m = {PGK_MTS_LONG, PGK_MTS_LONG, PGK_MTS_LONG, PGK_MTS_LONG,
               PGK_MTS_LONG, PGK_MTS_LONG, PGK_MTS_LONG};
  if (mts) m = *mts;
  const int32_t lv[7] = {m.bond, m.angl, m.tors, m.torsrb, m.vdw, m.ewk, m.nonb14};
  for (int k = 0; k < 7; k++)
    if (lv[k] != PGK_MTS_LONG && lv[k] != PGK_MTS_SHORT && lv[k] != 0) {
      pgk_set_error("pgk_md_respa: each term's level must be MTS_LONG, MTS_SHORT or 0");
      return PGK_EINVAL;
    }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  if (nstep_short > 1 && !c->d_md_fshort) {
    std::vector<double> z((size_t)3 * c->natom, 0.0);
    if ((rc = upload(&c->d_md_fshort, z))) return rc;
  }
  c->md_nshort = nstep_short;
  c->md_mts = m;
  return PGK_OK;
}

int pgk_md_langevin(pgk_ctx *c, const pgk_langevin_desc *lt) {
  if (!c || !c->md_ready) {
    pgk_set_error("pgk_md_langevin: no MD set-up");
    return PGK_EINVAL;
  }
  if (!lt) {
    c->md_lt_nreg = 0;
    c->md_lt_nstrm = 0;
    return PGK_OK;
  }
  const int nr = lt->nregion;
  const int ns = (lt->strm_z1 && lt->strm_z2 && lt->strm_vel) ? lt->nstream : 0;
  if (ns < 0 || ns > PGK_LT_MAXREG) {
    pgk_set_error("pgk_md_langevin: need 0 <= nstream <= PGK_LT_MAXREG");
    return PGK_EINVAL;
  }
  if (nr < 0 || nr > PGK_LT_MAXREG || (nr == 0 && ns == 0) ||
      (nr > 0 && (!lt->x1 || !lt->x2 || !lt->y1 || !lt->y2 || !lt->z1 || !lt->z2 ||
                  !lt->temperature || !lt->damp))) {
    pgk_set_error("pgk_md_langevin: need 1 <= nregion <= PGK_LT_MAXREG and every region array"
                  " (or nregion 0 with streaming slabs only)");
    return PGK_EINVAL;
  }
  for (int k = 0; k < nr; k++)
    if (!(lt->temperature[k] >= 0) || !(lt->damp[k] >= 0)) {
      pgk_set_error("pgk_md_langevin: temperatures and damping factors must be >= 0");
      return PGK_EINVAL;
    }
  const int n = c->natom;
  int nmol = 0, rc;
  for (int i = 0; i < n; i++) nmol = std::max(nmol, c->molid[i]);
  // registration units: molecules by centre of mass, or single atoms
  std::vector<int32_t> ustart(1, 0), uatom, ucom;
  std::vector<std::vector<int32_t>> mol(nmol);
  for (int i = 0; i < n; i++) mol[c->molid[i] - 1].push_back(i);
  for (int m = 0; m < nmol; m++) {
    if (mol[m].empty()) continue;
    if (!lt->mol_com || lt->mol_com[m]) {
      uatom.insert(uatom.end(), mol[m].begin(), mol[m].end());
      ustart.push_back((int32_t)uatom.size());
      ucom.push_back(1);
    } else {
      for (int32_t i : mol[m]) {
        uatom.push_back(i);
        ustart.push_back((int32_t)uatom.size());
        ucom.push_back(0);
      }
    }
  }
  if ((rc = set_device_ctx(c))) return rc;
  std::vector<double> zero((size_t)3 * n, 0.0);
  std::vector<int32_t> none(n, -1);
  if ((rc = upload(&c->d_lt_ustart, ustart)) || (rc = upload(&c->d_lt_uatom, uatom)) ||
      (rc = upload(&c->d_lt_ucom, ucom)) || (rc = upload(&c->d_lt_mass, c->md_mass)) ||
      (rc = upload(&c->d_lt_frand, zero)) || (rc = upload(&c->d_lt_reg, none)) ||
      (rc = upload(&c->d_lt_sreg, none)) || (rc = upload(&c->d_md_vstrm, zero)) ||
      (rc = upload(&c->d_lt_det, std::vector<double>(
                                     (size_t)pgk_grid_blocks(n, 256) * PGK_LT_MAXREG, 0.0))) ||
      (rc = upload(&c->d_lt_cnt, std::vector<int32_t>(PGK_LT_MAXREG, 0))))
    return rc;
  for (int k = 0; k < PGK_LT_MAXREG; k++) {
    const bool on = k < nr;
    c->md_lt_lo[0][k] = on ? lt->x1[k] : 0.0;
    c->md_lt_hi[0][k] = on ? lt->x2[k] : 0.0;
    c->md_lt_lo[1][k] = on ? lt->y1[k] : 0.0;
    c->md_lt_hi[1][k] = on ? lt->y2[k] : 0.0;
    c->md_lt_lo[2][k] = on ? lt->z1[k] : 0.0;
    c->md_lt_hi[2][k] = on ? lt->z2[k] : 0.0;
    c->md_lt_temp[k] = on ? lt->temperature[k] : 0.0;
    c->md_lt_damp[k] = on ? lt->damp[k] : 0.0;
  }
  for (int k = 0; k < PGK_LT_MAXREG; k++) {
    const bool on = k < ns;
    c->md_lt_sz[0][k] = on ? lt->strm_z1[k] : 0.0;
    c->md_lt_sz[1][k] = on ? lt->strm_z2[k] : 0.0;
    for (int d = 0; d < 3; d++) c->md_lt_sv[d][k] = on ? lt->strm_vel[3 * k + d] : 0.0;
  }
  c->md_lt_nstrm = ns;
  c->md_lt_nunit = (int)ucom.size();
  c->md_lt_seed = lt->seed;
  c->md_lt_draw = 0;
  c->md_lt_nreg = nr;
  return PGK_OK;
}

int pgk_md_langevin_state(pgk_ctx *c, double *det_ene_kin, int32_t *natom_lt) {
  if (!c || !c->md_lt_nreg) {
    pgk_set_error("pgk_md_langevin_state: no Langevin thermostat set up");
    return PGK_EINVAL;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  if ((rc = pgk_md_lt_count(c))) return rc;
  HIP_CHECK(hipStreamSynchronize(c->stream));
  const int nr = c->md_lt_nreg;
  if (det_ene_kin) {
    // per-block records, summed in block order
    const int nb = pgk_grid_blocks(c->natom, 256);
    std::vector<double> part((size_t)nb * PGK_LT_MAXREG);
    HIP_CHECK(hipMemcpy(part.data(), c->d_lt_det, sizeof(double) * part.size(),
                        hipMemcpyDeviceToHost));
    for (int k = 0; k < nr; k++) {
      double t = 0.0;
      for (int b = 0; b < nb; b++) t += part[(size_t)b * PGK_LT_MAXREG + k];
      det_ene_kin[k] = t;
    }
  }
  if (natom_lt)
    HIP_CHECK(hipMemcpy(natom_lt, c->d_lt_cnt, sizeof(int32_t) * nr, hipMemcpyDeviceToHost));
  return PGK_OK;
}

int pgk_md_heatf(pgk_ctx *c, const pgk_hf_desc *hf, int32_t interval) {
  if (!c || !c->md_ready) {
    pgk_set_error("pgk_md_heatf: no MD set-up");
    return PGK_EINVAL;
  }
  if (!hf || interval <= 0) {
    c->md_hf_interval = 0;
    return PGK_OK;
  }
  const int n = c->natom;
  if (hf->mode == PGK_HF_SURFACE || hf->nregion < 1 || !hf->z1 || !hf->z2 ||
      (hf->mode == PGK_HF_VOLUME && hf->nregion > PGK_HEATF_MAXREG)) {
    pgk_set_error("pgk_md_heatf: volume (up to " + std::to_string(PGK_HEATF_MAXREG) +
                  " regions) or bulk form needed");
    return PGK_EINVAL;
  }
  bool mole = false;
  for (int i = 0; hf->hftyp && i < n; i++) {
    if (hf->hftyp[i] != PGK_HFTYP_ATOM && hf->hftyp[i] != PGK_HFTYP_MOLE) {
      pgk_set_error("pgk_md_heatf: hftyp must be ATOM (1) or MOLE (2)");
      return PGK_EINVAL;
    }
    mole = mole || hf->hftyp[i] == PGK_HFTYP_MOLE;
  }
  int rc = set_device_ctx(c);
  if (rc) return rc;
  std::vector<double> m(n);
  for (int i = 0; i < n; i++) m[i] = c->md_mass[i] / PGK_MD_CONV;
  if ((rc = upload(&c->d_md_hfmass, m))) return rc;
  c->md_hf_z1.assign(hf->z1, hf->z1 + hf->nregion);
  c->md_hf_z2.assign(hf->z2, hf->z2 + hf->nregion);
  if (hf->hftyp) c->md_hf_typ.assign(hf->hftyp, hf->hftyp + n);
  else c->md_hf_typ.assign(n, PGK_HFTYP_ATOM);
  c->md_hf_mole = mole;
  c->md_hf_mode = hf->mode;
  c->md_hf_interval = interval;
  c->md_hf_rows = 0;
  return PGK_OK;
}

int pgk_md_heatf_series(pgk_ctx *c, double *out, int32_t max_rows, int32_t *nrows) {
  if (!c || !nrows || (max_rows > 0 && !out)) {
    pgk_set_error("pgk_md_heatf_series: null argument");
    return PGK_EINVAL;
  }
  const int rows = std::min<int>(c->md_hf_rows, std::max(max_rows, 0));
  *nrows = c->md_hf_rows;
  if (rows > 0) {
    int rc = set_device_ctx(c);
    if (rc) return rc;
    HIP_CHECK(hipStreamSynchronize(c->stream));
    HIP_CHECK(hipMemcpy(out, c->d_md_hfseries,
                        sizeof(double) * rows * c->md_hf_z1.size() * PGK_NHFJ * 3,
                        hipMemcpyDeviceToHost));
  }
  return PGK_OK;
}

int pgk_md_use_graphs(pgk_ctx *c, int32_t enable) {
  if (!c || enable < 0 || enable > 3) return PGK_EINVAL;
  c->md_graphs = (enable & 1) != 0;
  c->md_batch_graphs = (enable & 2) != 0;
  return PGK_OK;
}

int pgk_profile(pgk_ctx *c, int32_t enable, double *ms, int64_t *calls) {
  if (!c) return PGK_EINVAL;
  if (c->pn > 0) {
    HIP_CHECK(hipStreamSynchronize(c->stream));
    pgk_prof_collect(c);
  }
  if (ms)
    for (int k = 0; k < PGK_NPROF; k++) ms[k] = c->prof_ms[k];
  if (calls)
    for (int k = 0; k < PGK_NPROF; k++) calls[k] = c->prof_calls[k];
  if (enable >= 0) {
    c->prof = enable != 0;
    for (int k = 0; k < PGK_NPROF; k++) {
      c->prof_ms[k] = 0.0;
      c->prof_calls[k] = 0;
    }
  }
  return PGK_OK;
}

int pgk_profile_stages(pgk_ctx *c, uint32_t mask) {
  if (!c) return PGK_EINVAL;
  c->prof_mask = mask;
  return PGK_OK;
}

int pgk_tune(pgk_ctx *c, int32_t knob, int64_t value) {
  if (!c) {
    pgk_set_error("pgk_tune: null context");
    return PGK_EINVAL;
  }
  switch (knob) {
    case PGK_TUNE_PAIR_CHUNK_BELOW:
      if (value < 0) break;
      c->pair_chunk_below = value;
      return PGK_OK;
    case PGK_TUNE_PME_OVERLAP:
      if (value != 0 && value != 1) break;
      c->pme_overlap = value != 0;
      return PGK_OK;
    case PGK_TUNE_PAIR_CAP_BLOCKS:
      if (value < 0 || value > 16) break;
      c->pair_cap_default = (int)value;
      return PGK_OK;
    default:
      pgk_set_error("pgk_tune: unknown knob " + std::to_string(knob));
      return PGK_EINVAL;
  }
  pgk_set_error("pgk_tune: value out of range for knob " + std::to_string(knob));
  return PGK_EINVAL;
}

int pgk_list_info(pgk_ctx *c, int64_t *info) {
  if (!c || !info) return PGK_EINVAL;
  if (!c->have_list) {
    pgk_set_error("pgk_list_info: no list");
    return PGK_ESTATE;
  }
  int rc = set_device_ctx(c);
  if (rc || (rc = pgk_list_count_pairs(c))) return rc;
  info[0] = c->ncl;
  info[1] = c->npair_cl;
  info[2] = c->npairs;
  info[3] = PGK_CL * 4;   // pair slots per entry (8 i-slots x 4 j-slots)
  return PGK_OK;
}

int pgk_synchronize(pgk_ctx *c) {
  if (!c) return PGK_EINVAL;
  HIP_CHECK(hipStreamSynchronize(c->stream));
  pgk_prof_collect(c);
  return PGK_OK;
}

}  // extern "C"
