// k_list.hip - Verlet list rebuild on the GPU.
//
// Replaces mklist2a (mklist.F90:8-234) and mklist_cell (mklist_cell.F90:8-352).
// The listed pair set is exactly the reference's: unordered atom pairs with
// r^2 < rcut_book^2, where r is the FP64 minimum-image vector
// rij = x_i - x_j - L*dnint((x_i - x_j)/L) evaluated without fused
// multiply-add (mklist.F90:137-147), minus the exclusion list (mkexcl).
//
// Storage is MI355X-shaped rather than the reference's per-atom CSR: atoms are
// grouped into 8-atom clusters (sorted along z inside x-y columns), candidate
// cluster pairs are found with bounding boxes, and each kept cluster pair
// carries a 64-bit mask whose bit (ii + 8*jj) records the reference's decision
// for the pair (i-slot ii, j-slot jj).  The pair kernel evaluates exactly the
// masked pairs.
#include <hipcub/hipcub.hpp>

#include "pgk_internal.h"

namespace {

__device__ __forceinline__ double wrap_coord(double x, double L) {
  double w = x - L * floor(x / L);
  if (w >= L) w -= L;
  if (w < 0.0) w = 0.0;
  return w;
}

struct GridGeom {
  double L[3];
  double colw[2];
  double zbw;
  int ncol[2];
  int nzb;
};

__global__ void k_bin_atoms(const double *__restrict__ pos, int n, GridGeom g,
                            int32_t *__restrict__ bin_count,
                            int32_t *__restrict__ atom_bin,
                            int32_t *__restrict__ atom_slot) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double wx = wrap_coord(pos[3 * i + 0], g.L[0]);
  double wy = wrap_coord(pos[3 * i + 1], g.L[1]);
  double wz = wrap_coord(pos[3 * i + 2], g.L[2]);
  int cx = max(0, min((int)(wx / g.colw[0]), g.ncol[0] - 1));
  int cy = max(0, min((int)(wy / g.colw[1]), g.ncol[1] - 1));
  int zb = max(0, min((int)(wz / g.zbw), g.nzb - 1));
  int bin = (cy * g.ncol[0] + cx) * g.nzb + zb;
  atom_bin[i] = bin;
  atom_slot[i] = atomicAdd(&bin_count[bin], 1);
}

__global__ void k_scatter(int n, const int32_t *__restrict__ atom_bin,
                          const int32_t *__restrict__ atom_slot,
                          const int32_t *__restrict__ bin_start,
                          int32_t *__restrict__ sorted) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sorted[bin_start[atom_bin[i]] + atom_slot[i]] = i;
}

// deterministic order inside a bin: ascending atom index (bins hold ~1 atom)
__global__ void k_sort_bins(int nbin, const int32_t *__restrict__ bin_start,
                            int32_t *__restrict__ sorted) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbin) return;
  int s = bin_start[b], e = bin_start[b + 1];
  for (int k = s + 1; k < e; k++) {
    int v = sorted[k];
    int m = k - 1;
    while (m >= s && sorted[m] > v) {
      sorted[m + 1] = sorted[m];
      m--;
    }
    sorted[m + 1] = v;
  }
}

__global__ void k_col_counts(int ncolumn, int nzb, const int32_t *__restrict__ bin_start,
                             int32_t *__restrict__ col_ncl) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncolumn) return;
  int cnt = bin_start[(c + 1) * nzb] - bin_start[c * nzb];
  col_ncl[c] = (cnt + PGK_CL - 1) / PGK_CL;
}

// one block per column
__global__ void k_fill_clusters(int nzb, const int32_t *__restrict__ bin_start,
                                const int32_t *__restrict__ sorted,
                                const int32_t *__restrict__ col_ncl,
                                const int32_t *__restrict__ col_clstart,
                                int32_t *__restrict__ cl_atom,
                                int32_t *__restrict__ cl_col,
                                int32_t *__restrict__ slot_of) {
  int c = blockIdx.x;
  int s0 = bin_start[c * nzb];
  int cnt = bin_start[(c + 1) * nzb] - s0;
  int nslot = col_ncl[c] * PGK_CL;
  int base = col_clstart[c] * PGK_CL;
  for (int k = threadIdx.x; k < nslot; k += blockDim.x) {
    int a = (k < cnt) ? sorted[s0 + k] : -1;
    cl_atom[base + k] = a;
    if (a >= 0) slot_of[a] = base + k;
    if ((k & (PGK_CL - 1)) == 0) cl_col[col_clstart[c] + k / PGK_CL] = c;
  }
}

__global__ void k_bbox(int ncl, const double *__restrict__ pos,
                       const int32_t *__restrict__ cl_atom, GridGeom g,
                       float *__restrict__ bbox) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncl) return;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int k = 0; k < PGK_CL; k++) {
    int a = cl_atom[c * PGK_CL + k];
    if (a < 0) continue;
    for (int d = 0; d < 3; d++) {
      double w = wrap_coord(pos[3 * a + d], g.L[d]);
      lo[d] = fmin(lo[d], w);
      hi[d] = fmax(hi[d], w);
    }
  }
  for (int d = 0; d < 3; d++) {
    bbox[c * 8 + d] = (float)lo[d] - 1e-3f;
    bbox[c * 8 + 4 + d] = (float)hi[d] + 1e-3f;
  }
  bbox[c * 8 + 3] = 0.f;
  bbox[c * 8 + 7] = 0.f;
}

// minimal periodic gap between intervals [a0,a1] and [b0,b1] in a box of length L
__device__ __forceinline__ float pgap(float a0, float a1, float b0, float b1, float L) {
  float g = 3.0e38f;
  for (int s = -1; s <= 1; s++) {
    float bb0 = b0 + s * L, bb1 = b1 + s * L;
    float t = fmaxf(0.f, fmaxf(bb0 - a1, a0 - bb1));
    g = fminf(g, t);
  }
  return g;
}

// exact reference pair test (no contraction), 0-based atoms
__device__ __forceinline__ bool ref_pair_in(const double *__restrict__ pos, int a, int b,
                                            const double *L, const double *Linv,
                                            double sqcut) {
#pragma clang fp contract(off)
  double r[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    r[d] = pos[3 * a + d] - pos[3 * b + d];
    r[d] = r[d] - L[d] * round(r[d] * Linv[d]);
  }
  double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  return r2 < sqcut;
}

__device__ __forceinline__ bool excluded(const int32_t *__restrict__ xi,
                                         const int32_t *__restrict__ xl, int a, int b) {
  int i = min(a, b), j = max(a, b) + 1;   // j 1-based
  int s = xi[i] - 1, e = xi[i + 1] - 1;
  for (int k = s; k < e; k++)
    if (xl[k] == j) return true;
  return false;
}

struct ListArgs {
  const double *pos;
  const int32_t *cl_atom, *cl_col, *col_ncl, *col_clstart;
  const float *bbox;
  const int32_t *xidx, *xlst;
  int32_t *tmp_cj;
  uint64_t *tmp_mask;
  int32_t *ci_count;
  int32_t *overflow;
  int ncl;
  int ncol[2];
  int rng[2];
  double L[3], Linv[3];
  float Lf[3];
  double sqcut;
  float rcbf2;
};

// one wavefront per i-cluster
__global__ __launch_bounds__(256) void k_find_pairs(ListArgs a) {
  __shared__ int32_t s_cand[4][PGK_MAXJ];
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int ci = blockIdx.x * 4 + wave;
  const bool active = ci < a.ncl;
  int32_t *cand = s_cand[wave];
  bool ovf = false;
  int ncand = 0;
  if (active) {
    float ilo[3], ihi[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      ilo[d] = a.bbox[ci * 8 + d];
      ihi[d] = a.bbox[ci * 8 + 4 + d];
    }
    int col = a.cl_col[ci];
    int cx = col % a.ncol[0], cy = col / a.ncol[0];
    int nx = min(2 * a.rng[0] + 1, a.ncol[0]);
    int ny = min(2 * a.rng[1] + 1, a.ncol[1]);
    int x0 = (nx == a.ncol[0]) ? 0 : cx - a.rng[0];
    int y0 = (ny == a.ncol[1]) ? 0 : cy - a.rng[1];
    // phase A: bounding-box filter, candidates compacted into LDS; the running
    // count is wave-uniform (derived from ballots) and kept in a register
    for (int iy = 0; iy < ny; iy++) {
      int yy = ((y0 + iy) % a.ncol[1] + a.ncol[1]) % a.ncol[1];
      for (int ix = 0; ix < nx; ix++) {
        int xx = ((x0 + ix) % a.ncol[0] + a.ncol[0]) % a.ncol[0];
        int cc = yy * a.ncol[0] + xx;
        int c0 = a.col_clstart[cc], cn = a.col_ncl[cc];
        for (int k0 = 0; k0 < cn; k0 += 64) {
          int cj = c0 + k0 + lane;
          bool pass = false;
          if (k0 + lane < cn && cj >= ci) {
            float g2 = 0.f;
#pragma unroll
            for (int d = 0; d < 3; d++) {
              float gd = pgap(ilo[d], ihi[d], a.bbox[cj * 8 + d], a.bbox[cj * 8 + 4 + d],
                              a.Lf[d]);
              g2 += gd * gd;
            }
            pass = g2 < a.rcbf2;
          }
          uint64_t bal = __ballot(pass);
          int off = __popcll(bal & ((1ull << lane) - 1ull));
          if (pass) {
            if (ncand + off < PGK_MAXJ) cand[ncand + off] = cj;
            else ovf = true;
          }
          ncand = min(ncand + (int)__popcll(bal), PGK_MAXJ);
        }
      }
    }
  }
  __syncthreads();
  if (active) {
    // phase B: exact per-pair masks, one candidate per iteration, lane = pair
    int ii = lane & 7, jj = lane >> 3;
    int ai = a.cl_atom[ci * PGK_CL + ii];
    int count = 0;
    for (int k = 0; k < ncand; k++) {
      int cj = cand[k];
      int aj = a.cl_atom[cj * PGK_CL + jj];
      bool in = false;
      if (ai >= 0 && aj >= 0 && (cj != ci || ii < jj)) {
        int lo = min(ai, aj), hi = max(ai, aj);
        in = ref_pair_in(a.pos, lo, hi, a.L, a.Linv, a.sqcut);
        if (in) in = !excluded(a.xidx, a.xlst, lo, hi);
      }
      uint64_t m = __ballot(in);
      if (m) {
        if (lane == 0) {
          if (count < PGK_MAXJ) {
            a.tmp_cj[(size_t)ci * PGK_MAXJ + count] = cj;
            a.tmp_mask[(size_t)ci * PGK_MAXJ + count] = m;
          } else {
            ovf = true;
          }
        }
        count++;
      }
    }
    if (lane == 0) a.ci_count[ci] = min(count, PGK_MAXJ);
    if (__ballot(ovf) && lane == 0) atomicOr(a.overflow, 1);
  }
}

__global__ void k_compact_pairs(int ncl, const int32_t *__restrict__ ci_count,
                                const int32_t *__restrict__ ci_start,
                                const int32_t *__restrict__ tmp_cj,
                                const uint64_t *__restrict__ tmp_mask,
                                int32_t *__restrict__ cj, uint64_t *__restrict__ mask,
                                unsigned long long *__restrict__ npairs) {
  int ci = blockIdx.x;
  int n = ci_count[ci], s = ci_start[ci];
  unsigned long long cnt = 0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    uint64_t m = tmp_mask[(size_t)ci * PGK_MAXJ + k];
    cj[s + k] = tmp_cj[(size_t)ci * PGK_MAXJ + k];
    mask[s + k] = m;
    cnt += __popcll(m);
  }
  if (cnt) atomicAdd(npairs, cnt);
}

template <class T>
int dev_alloc(T **p, size_t n) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    pgk_set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return PGK_ENOMEM;
  }
  return PGK_OK;
}

}  // namespace

int pgk_list_build(pgk_ctx *c, double rcut_book) {
  const int n = c->natom;
  hipStream_t st = c->stream;
  if (!c->have_pos) {
    pgk_set_error("pgk_build_list: positions not set");
    return PGK_ESTATE;
  }
  GridGeom g;
  for (int d = 0; d < 3; d++) g.L[d] = c->box[d];
  double vol = g.L[0] * g.L[1] * g.L[2];
  double rho = (double)n / vol;
  double w = cbrt((double)PGK_CL / fmax(rho, 1e-6));
  for (int d = 0; d < 2; d++) {
    g.ncol[d] = (int)fmax(1.0, floor(g.L[d] / w));
    g.colw[d] = g.L[d] / g.ncol[d];
  }
  g.nzb = (int)fmax(1.0, floor(g.L[2] / 0.25));
  if (g.nzb > 8192) g.nzb = 8192;
  g.zbw = g.L[2] / g.nzb;
  int ncolumn = g.ncol[0] * g.ncol[1];
  int nbin = ncolumn * g.nzb;
  int rc;
  if (nbin + 1 > c->nbin_cap) {
    if ((rc = dev_alloc(&c->d_bin_count, nbin + 1)) || (rc = dev_alloc(&c->d_bin_start, nbin + 1)))
      return rc;
    c->nbin_cap = nbin + 1;
  }
  if (ncolumn + 1 > c->ncolumn_cap) {
    if ((rc = dev_alloc(&c->d_col_ncl, ncolumn + 1)) ||
        (rc = dev_alloc(&c->d_col_clstart, ncolumn + 1)))
      return rc;
    c->ncolumn_cap = ncolumn + 1;
  }
  if (!c->d_atom_bin) {
    if ((rc = dev_alloc(&c->d_atom_bin, n)) || (rc = dev_alloc(&c->d_atom_slot, n)) ||
        (rc = dev_alloc(&c->d_sorted, n)) || (rc = dev_alloc(&c->d_slot_of, n)) ||
        (rc = dev_alloc(&c->d_overflow, 4)))
      return rc;
  }
  // scan scratch sized for the larger of the two scans
  size_t need = 0, t1 = 0, t2 = 0;
  hipcub::DeviceScan::ExclusiveSum(nullptr, t1, c->d_bin_count, c->d_bin_start, nbin + 1, st);
  int ncl_max = n / PGK_CL + ncolumn + 1;
  hipcub::DeviceScan::ExclusiveSum(nullptr, t2, c->d_bin_count, c->d_bin_start, ncl_max + 1, st);
  need = t1 > t2 ? t1 : t2;
  if (need > c->scan_tmp_bytes) {
    if (c->d_scan_tmp) (void)hipFree(c->d_scan_tmp);
    c->d_scan_tmp = nullptr;
    HIP_CHECK(hipMalloc(&c->d_scan_tmp, need));
    c->scan_tmp_bytes = need;
  }
  HIP_CHECK(hipMemsetAsync(c->d_bin_count, 0, sizeof(int32_t) * (nbin + 1), st));
  HIP_CHECK(hipMemsetAsync(c->d_overflow, 0, sizeof(int32_t), st));
  k_bin_atoms<<<pgk_grid_blocks(n, 256), 256, 0, st>>>(c->d_pos, n, g, c->d_bin_count,
                                                        c->d_atom_bin, c->d_atom_slot);
  HIP_CHECK(hipGetLastError());
  size_t tb = c->scan_tmp_bytes;
  HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(c->d_scan_tmp, tb, c->d_bin_count, c->d_bin_start,
                                             nbin + 1, st));
  k_scatter<<<pgk_grid_blocks(n, 256), 256, 0, st>>>(n, c->d_atom_bin, c->d_atom_slot,
                                                      c->d_bin_start, c->d_sorted);
  k_sort_bins<<<pgk_grid_blocks(nbin, 256), 256, 0, st>>>(nbin, c->d_bin_start, c->d_sorted);
  k_col_counts<<<pgk_grid_blocks(ncolumn, 256), 256, 0, st>>>(ncolumn, g.nzb, c->d_bin_start,
                                                              c->d_col_ncl);
  HIP_CHECK(hipMemsetAsync(c->d_col_ncl + ncolumn, 0, sizeof(int32_t), st));
  tb = c->scan_tmp_bytes;
  HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(c->d_scan_tmp, tb, c->d_col_ncl, c->d_col_clstart,
                                             ncolumn + 1, st));
  int32_t ncl = 0;
  HIP_CHECK(hipMemcpyAsync(&ncl, c->d_col_clstart + ncolumn, sizeof(int32_t),
                           hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  if (ncl > c->ncl_cap) {
    size_t ns = (size_t)ncl * PGK_CL;
    if ((rc = dev_alloc(&c->d_cl_atom, ns)) || (rc = dev_alloc(&c->d_cl_col, ncl)) ||
        (rc = dev_alloc(&c->d_bbox, (size_t)ncl * 8)) ||
        (rc = dev_alloc(&c->d_tmp_cj, (size_t)ncl * PGK_MAXJ)) ||
        (rc = dev_alloc(&c->d_tmp_mask, (size_t)ncl * PGK_MAXJ)) ||
        (rc = dev_alloc(&c->d_ci_count, (size_t)ncl + 1)) ||
        (rc = dev_alloc(&c->d_ci_start, (size_t)ncl + 1)) ||
        (rc = dev_alloc((int32_t **)&c->d_ct, ns)))
      return rc;
    size_t xqb = (c->precision == PGK_PREC_FP64) ? 32 : 16;
    if (c->d_xq) (void)hipFree(c->d_xq);
    c->d_xq = nullptr;
    HIP_CHECK(hipMalloc(&c->d_xq, ns * xqb));
    size_t fb = (c->precision == PGK_PREC_FP32) ? 4 : 8;
    if (c->d_fcl) (void)hipFree(c->d_fcl);
    if (c->d_fcl_c) (void)hipFree(c->d_fcl_c);
    c->d_fcl = c->d_fcl_c = nullptr;
    HIP_CHECK(hipMalloc(&c->d_fcl, ns * 3 * fb));
    HIP_CHECK(hipMalloc(&c->d_fcl_c, ns * 3 * fb));
    c->ncl_cap = ncl;
    c->acc_dirty = true;   // new accumulators start zeroed by the next call
  }
  c->ncl = ncl;
  c->ncol[0] = g.ncol[0];
  c->ncol[1] = g.ncol[1];
  c->nzb = g.nzb;
  k_fill_clusters<<<ncolumn, 64, 0, st>>>(g.nzb, c->d_bin_start, c->d_sorted, c->d_col_ncl,
                                          c->d_col_clstart, c->d_cl_atom, c->d_cl_col,
                                          c->d_slot_of);
  k_bbox<<<pgk_grid_blocks(ncl, 128), 128, 0, st>>>(ncl, c->d_pos, c->d_cl_atom, g, c->d_bbox);
  HIP_CHECK(hipGetLastError());

  ListArgs a;
  a.pos = c->d_pos;
  a.cl_atom = c->d_cl_atom;
  a.cl_col = c->d_cl_col;
  a.col_ncl = c->d_col_ncl;
  a.col_clstart = c->d_col_clstart;
  a.bbox = c->d_bbox;
  a.xidx = c->d_excl_index;
  a.xlst = c->d_excl_list;
  a.tmp_cj = c->d_tmp_cj;
  a.tmp_mask = c->d_tmp_mask;
  a.ci_count = c->d_ci_count;
  a.overflow = c->d_overflow;
  a.ncl = ncl;
  for (int d = 0; d < 2; d++) {
    a.ncol[d] = g.ncol[d];
    a.rng[d] = (int)ceil(rcut_book / g.colw[d]);
  }
  for (int d = 0; d < 3; d++) {
    a.L[d] = g.L[d];
    a.Linv[d] = 1.0 / g.L[d];
    a.Lf[d] = (float)g.L[d];
  }
  a.sqcut = rcut_book * rcut_book;
  float rf = (float)rcut_book + 1e-3f;
  a.rcbf2 = rf * rf;
  k_find_pairs<<<pgk_grid_blocks(ncl, 4), 256, 0, st>>>(a);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemsetAsync(c->d_ci_count + ncl, 0, sizeof(int32_t), st));
  tb = c->scan_tmp_bytes;
  HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(c->d_scan_tmp, tb, c->d_ci_count, c->d_ci_start,
                                             ncl + 1, st));
  int32_t tot = 0, ovf = 0;
  HIP_CHECK(hipMemcpyAsync(&tot, c->d_ci_start + ncl, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipMemcpyAsync(&ovf, c->d_overflow, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  if (ovf) {
    pgk_set_error("pgk_build_list: more than 2048 neighbour clusters for one cluster "
                  "(rcut_book too large for the cluster list)");
    return PGK_EOVERFLOW;
  }
  if (tot > c->cap_pair_cl) {
    int64_t cap = (int64_t)(tot * 1.25) + 1024;
    if ((rc = dev_alloc(&c->d_cj, cap)) || (rc = dev_alloc(&c->d_mask, cap))) return rc;
    c->cap_pair_cl = cap;
  }
  c->npair_cl = tot;
  unsigned long long *d_np = reinterpret_cast<unsigned long long *>(c->d_sum);
  HIP_CHECK(hipMemsetAsync(d_np, 0, sizeof(unsigned long long), st));
  k_compact_pairs<<<ncl, 64, 0, st>>>(ncl, c->d_ci_count, c->d_ci_start, c->d_tmp_cj,
                                      c->d_tmp_mask, c->d_cj, c->d_mask, d_np);
  HIP_CHECK(hipGetLastError());
  unsigned long long np = 0;
  HIP_CHECK(hipMemcpyAsync(&np, d_np, sizeof(np), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  c->npairs = (int64_t)np;
  c->have_list = true;
  for (int d = 0; d < 3; d++) c->list_box[d] = c->box[d];
  return PGK_OK;
}
