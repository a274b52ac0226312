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
// cluster pairs are found with bounding boxes, and every half j-cluster (j
// slots 0-3 or 4-7) with a listed pair becomes an entry: half-cluster index
// 2 cj + h and a 32-bit mask whose bit (ii + 8 jq) records the reference's
// decision for the pair (i-slot ii, j-slot 4 h + jq).  The pair kernel
// evaluates exactly the masked pairs, two entries per wavefront pass.
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

// bounding boxes of the clusters (wrapped frame) and the slots' FP64
// positions with the atom index in .w (-1 for padding), for the pair search
__global__ void k_bbox(const int32_t *__restrict__ ncl, const double *__restrict__ pos,
                       const int32_t *__restrict__ cl_atom, GridGeom g,
                       float *__restrict__ bbox, double4 *__restrict__ posc) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= *ncl) return;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int k = 0; k < PGK_CL; k++) {
    int a = cl_atom[c * PGK_CL + k];
    posc[c * PGK_CL + k] = (a >= 0) ? make_double4(pos[3 * a], pos[3 * a + 1], pos[3 * a + 2],
                                                   (double)a)
                                    : make_double4(0.0, 0.0, 0.0, -1.0);
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

// the reference's pair test (mklist.F90:134-145, FP64, no contraction) on two
// packed positions: the minimum-image r^2 does not depend on which atom comes
// first (IEEE subtraction and round() are odd functions), so the slot order
// can be used
__device__ __forceinline__ bool ref_pair_in4(const double4 &p, const double4 &q, const double *L,
                                             const double *Linv, double sqcut) {
#pragma clang fp contract(off)
  double r[3] = {p.x - q.x, p.y - q.y, p.z - q.z};
#pragma unroll
  for (int d = 0; d < 3; d++) r[d] = r[d] - L[d] * round(r[d] * Linv[d]);
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

#define PGK_XREG 8   // exclusion partners of the i-atom kept in registers
#define PGK_UNROLL 4 // candidates whose j-slots are loaded together

struct ListArgs {
  const double4 *posc;
  const int32_t *xs_start, *xs_j;   // symmetric CSR of the exclusions j > i, 0-based
  const int32_t *cl_col, *col_ncl, *col_clstart;
  const float *bbox;
  const int32_t *xidx, *xlst;
  int32_t *tmp_cj;
  uint32_t *tmp_mask;
  int32_t *ci_count;
  int32_t *overflow;
  int ncolumn, ncl_max;
  int ncol[2], rng[2];
  double L[3], Linv[3];
  float Lf[3], colwf[2];
  float zwin;      // half-height of a column's z window around the i-cluster box
  double sqcut;
  float rcbf2;
};

// first k in [0, n) with key(k) > thr if gt, else first k with key(k) >= thr;
// keys are the clusters' z bounds, nondecreasing up to one z bin
__device__ __forceinline__ int zsearch(const float *__restrict__ bbox, int c0, int n, int off,
                                       float thr, bool gt) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const float v = bbox[(c0 + mid) * 8 + off];
    if (gt ? (v > thr) : (v >= thr)) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// One wavefront per i-cluster.
// Phase A walks the x-y columns within reach, at or after the i-cluster's own
// column (cluster indices are column-major, and only cj >= ci is kept).  For
// each column one lane finds by binary search the clusters whose z range can
// come within rcut_book of the i-cluster's box (a circular window when the
// reach crosses the periodic z boundary); the windows of up to 64 columns are
// then tested 64 clusters at a time with the bounding-box gap, and survivors
// are compacted into LDS.
// Phase B decides every atom pair of every surviving cluster pair with the
// reference's exact FP64 test, lane = ii + 8*jj, PGK_UNROLL candidates per
// iteration, and writes (cj, 64-bit mask) for nonzero masks.
// NW waves per block: 4, or 1 for small systems, whose few clusters then
// spread over all CUs instead of a quarter as many blocks
template <int NW>
__global__ __launch_bounds__(64 * NW) void k_find_pairs(ListArgs a) {
  __shared__ int32_t s_cand[NW][PGK_MAXJ];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ci = blockIdx.x * NW + wave;
  const int ncl = a.col_clstart[a.ncolumn];
  const bool active = ci < ncl;
  int32_t *cand = s_cand[wave];
  bool ovf = false;
  int ncand = 0;
  if (active) {
    const float4 blo = *reinterpret_cast<const float4 *>(a.bbox + ci * 8);
    const float4 bhi = *reinterpret_cast<const float4 *>(a.bbox + ci * 8 + 4);
    const float ilo[3] = {blo.x, blo.y, blo.z}, ihi[3] = {bhi.x, bhi.y, bhi.z};
    const int col = a.cl_col[ci];
    const int cx = col % a.ncol[0], cy = col / a.ncol[0];
    const int nx = min(2 * a.rng[0] + 1, a.ncol[0]);
    const int ny = min(2 * a.rng[1] + 1, a.ncol[1]);
    const int x0 = (nx == a.ncol[0]) ? 0 : cx - a.rng[0];
    const int y0 = (ny == a.ncol[1]) ? 0 : cy - a.rng[1];
    // z window: 0 whole column, 1 plain [zs, ze], 2 circular (suffix from
    // zs, prefix up to ze, both already mapped into [0, L))
    float zs = ilo[2] - a.zwin, ze = ihi[2] + a.zwin;
    int zmode = 1;
    if (ze - zs >= a.Lf[2]) {
      zmode = 0;
    } else if (zs < 0.f) {
      zmode = 2;
      zs += a.Lf[2];
    } else if (ze > a.Lf[2]) {
      zmode = 2;
      ze -= a.Lf[2];
    }
    const int nwin = nx * ny;
    for (int w0 = 0; w0 < nwin; w0 += 64) {
      const int w = w0 + lane;
      int c0 = 0, cn = 0, start = 0, len = 0;
      if (w < nwin) {
        const int ix = w % nx, iy = w / nx;
        const int xx = ((x0 + ix) % a.ncol[0] + a.ncol[0]) % a.ncol[0];
        const int yy = ((y0 + iy) % a.ncol[1] + a.ncol[1]) % a.ncol[1];
        const int cc = yy * a.ncol[0] + xx;
        if (cc >= col) {
          const float gx = pgap(ilo[0], ihi[0], xx * a.colwf[0] - 0.01f,
                                (xx + 1) * a.colwf[0] + 0.01f, a.Lf[0]);
          const float gy = pgap(ilo[1], ihi[1], yy * a.colwf[1] - 0.01f,
                                (yy + 1) * a.colwf[1] + 0.01f, a.Lf[1]);
          if (gx * gx + gy * gy < a.rcbf2) {
            c0 = a.col_clstart[cc];
            cn = a.col_ncl[cc];
            if (zmode == 0) {
              len = cn;
            } else if (zmode == 1) {
              start = zsearch(a.bbox, c0, cn, 6, zs, false);   // hi_z >= zs
              const int end = zsearch(a.bbox, c0, cn, 2, ze, true);  // lo_z > ze
              len = max(0, end - start);
            } else {
              start = zsearch(a.bbox, c0, cn, 6, zs, false);
              const int pre = min(zsearch(a.bbox, c0, cn, 2, ze, true), start);
              len = (cn - start) + pre;
            }
          }
        }
      }
      int incl = len;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
      }
      const int off = incl - len;
      const int total = __shfl(incl, 63);
      for (int t0 = 0; t0 < total; t0 += 64) {
        // every lane runs the shuffles (a shuffle from an inactive lane reads
        // nothing useful); lanes past the end repeat the last element
        const int t = t0 + lane;
        const int tt = min(t, total - 1);
        int lo = 0;   // the window holding tt: the last lane whose offset is <= tt
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1) {
          const int o2 = __shfl(off, (lo + step) & 63);
          if (lo + step < 64 && o2 <= tt) lo += step;
        }
        const int wc0 = __shfl(c0, lo), wcn = __shfl(cn, lo), wst = __shfl(start, lo);
        int k = wst + (tt - __shfl(off, lo));
        if (k >= wcn) k -= wcn;
        const int cj = wc0 + k;
        bool pass = false;
        if (t < total && cj >= ci) {
          const float4 jlo = *reinterpret_cast<const float4 *>(a.bbox + cj * 8);
          const float4 jhi = *reinterpret_cast<const float4 *>(a.bbox + cj * 8 + 4);
          const float g0 = pgap(ilo[0], ihi[0], jlo.x, jhi.x, a.Lf[0]);
          const float g1 = pgap(ilo[1], ihi[1], jlo.y, jhi.y, a.Lf[1]);
          const float g2 = pgap(ilo[2], ihi[2], jlo.z, jhi.z, a.Lf[2]);
          pass = g0 * g0 + g1 * g1 + g2 * g2 < a.rcbf2;
        }
        const uint64_t bal = __ballot(pass);
        const int pos = __popcll(bal & ((1ull << lane) - 1ull));
        if (pass) {
          if (ncand + pos < PGK_MAXJ) cand[ncand + pos] = cj;
          else ovf = true;
        }
        ncand = min(ncand + (int)__popcll(bal), PGK_MAXJ);
      }
    }
  }
  __syncthreads();
  if (active) {
    const int ii = lane & 7, jj = lane >> 3;
    const double4 pi4 = a.posc[ci * PGK_CL + ii];
    const int ai = (int)pi4.w;
    int xp[PGK_XREG];
    bool xmem = false;
    {
      int s0 = 0, nxp = 0;
      if (ai >= 0) {
        s0 = a.xs_start[ai];
        nxp = a.xs_start[ai + 1] - s0;
      }
      xmem = nxp > PGK_XREG;
#pragma unroll
      for (int q = 0; q < PGK_XREG; q++) xp[q] = (q < nxp) ? a.xs_j[s0 + q] : -1;
    }
    int count = 0;
    for (int k0 = 0; k0 < ncand; k0 += PGK_UNROLL) {
      int cjv[PGK_UNROLL];
      double4 pj[PGK_UNROLL];
#pragma unroll
      for (int u = 0; u < PGK_UNROLL; u++) {
        cjv[u] = cand[min(k0 + u, ncand - 1)];
        pj[u] = a.posc[cjv[u] * PGK_CL + jj];
      }
#pragma unroll
      for (int u = 0; u < PGK_UNROLL; u++) {
        if (k0 + u >= ncand) break;
        const int cj = cjv[u];
        const int aj = (int)pj[u].w;
        bool in = false;
        if (ai >= 0 && aj >= 0 && (cj != ci || ii < jj)) {
          in = ref_pair_in4(pi4, pj[u], a.L, a.Linv, a.sqcut);
          if (in) {
            bool ex = false;
#pragma unroll
            for (int q = 0; q < PGK_XREG; q++) ex = ex || (xp[q] == aj);
            if (xmem) ex = excluded(a.xidx, a.xlst, ai, aj);
            in = !ex;
          }
        }
        const uint64_t m = __ballot(in);
        // one entry per half j-cluster with a listed pair: slots jj = 0..3
        // (mask bits 0..31) and jj = 4..7 (bits 32..63)
#pragma unroll
        for (int hh = 0; hh < 2; hh++) {
          const uint32_t mh = (uint32_t)(m >> (32 * hh));
          if (mh) {
            if (lane == 0) {
              if (count < PGK_MAXJ) {
                a.tmp_cj[(size_t)ci * PGK_MAXJ + count] = 2 * cj + hh;
                a.tmp_mask[(size_t)ci * PGK_MAXJ + count] = mh;
              } else {
                ovf = true;
              }
            }
            count++;
          }
        }
      }
    }
    if (lane == 0) a.ci_count[ci] = min(count, PGK_MAXJ);
    if (__ballot(ovf) && lane == 0) atomicOr(a.overflow, 1);
  } else if (ci <= a.ncl_max && lane == 0) {
    a.ci_count[ci] = 0;   // slots past the real cluster count scan as empty
  }
}

// copy each cluster's entries to its place in the compact list; a list
// buffer that is too small is reported (bit 2 of *overflow) and not written
__global__ void k_compact_pairs(int ncl_max, const int32_t *__restrict__ ci_count,
                                const int32_t *__restrict__ ci_start,
                                const int32_t *__restrict__ tmp_cj,
                                const uint32_t *__restrict__ tmp_mask,
                                int32_t *__restrict__ cj, uint32_t *__restrict__ mask,
                                int64_t cap, int32_t *__restrict__ overflow) {
  const int ci = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (ci >= ncl_max) return;
  if (ci == 0 && lane == 0 && ci_start[ncl_max] > cap) atomicOr(overflow, 2);
  const int n = ci_count[ci], s = ci_start[ci];
  if ((int64_t)s + n > cap) return;
  for (int k = lane; k < n; k += 64) {
    cj[s + k] = tmp_cj[(size_t)ci * PGK_MAXJ + k];
    mask[s + k] = tmp_mask[(size_t)ci * PGK_MAXJ + k];
  }
}

// Work items of the pair kernel's chunked form: chunk k (entries 64 k ..
// 64 k + 63 of its list) of cluster ci, coded ci * 32 + k (a cluster holds at
// most PGK_MAXJ = 2048 entries, 32 chunks).
__global__ void k_chunk_count(int ncl_max, const int32_t *__restrict__ ci_count,
                              int32_t *__restrict__ wcount) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci <= ncl_max) wcount[ci] = (ci < ncl_max) ? (ci_count[ci] + 63) / 64 : 0;
}
__global__ void k_chunk_fill(int ncl_max, const int32_t *__restrict__ wcount,
                             const int32_t *__restrict__ wstart, int32_t *__restrict__ witem,
                             int64_t cap) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= ncl_max) return;
  for (int k = 0; k < wcount[ci]; k++) {
    const int64_t w = (int64_t)wstart[ci] + k;
    if (w < cap) witem[w] = ci * 32 + k;
  }
}

// cluster count, overflow flags, entry count and work-item count in one
// record for the host
__global__ void k_list_meta(const int32_t *__restrict__ ncl, const int32_t *__restrict__ ovf,
                            const int32_t *__restrict__ tot, const int32_t *__restrict__ wtot,
                            int32_t *__restrict__ meta) {
  if (threadIdx.x == 0) {
    meta[0] = *ncl;
    meta[1] = *ovf;
    meta[2] = *tot;
    meta[3] = *wtot;
  }
}

// number of listed atom pairs (bits set over all masks)
__global__ void k_count_pairs(int64_t n, const uint32_t *__restrict__ mask,
                              unsigned long long *__restrict__ out) {
  __shared__ unsigned long long s_part[4];
  unsigned long long cnt = 0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    cnt += __popc(mask[k]);
  for (int o = 32; o >= 1; o >>= 1) cnt += __shfl_down(cnt, o);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, s_part[0] + s_part[1] + s_part[2] + s_part[3]);
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
  // every column rounds its atoms up to whole clusters, so the cluster count
  // is at most n/8 + ncolumn; buffers are sized for that bound and the
  // kernels read the real count from the column scan, so nothing waits for
  // the host until the single read-back at the end
  const int ncl_max = n / PGK_CL + ncolumn + 1;
  size_t need = 0, t1 = 0, t2 = 0;
  HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, t1, c->d_bin_count, c->d_bin_start,
                                             nbin + 1, st));
  HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, t2, c->d_bin_count, c->d_bin_start,
                                             ncl_max + 1, st));
  need = t1 > t2 ? t1 : t2;
  if (need > c->scan_tmp_bytes) {
    if (c->d_scan_tmp) (void)hipFree(c->d_scan_tmp);
    c->d_scan_tmp = nullptr;
    HIP_CHECK(hipMalloc(&c->d_scan_tmp, need));
    c->scan_tmp_bytes = need;
  }
  if (ncl_max > c->ncl_cap) {
    size_t ns = (size_t)ncl_max * PGK_CL;
    if ((rc = dev_alloc(&c->d_cl_atom, ns)) || (rc = dev_alloc(&c->d_cl_col, ncl_max)) ||
        (rc = dev_alloc(&c->d_bbox, (size_t)ncl_max * 8)) ||
        (rc = dev_alloc(&c->d_posc, ns)) ||
        (rc = dev_alloc(&c->d_tmp_cj, (size_t)ncl_max * PGK_MAXJ)) ||
        (rc = dev_alloc(&c->d_tmp_mask, (size_t)ncl_max * PGK_MAXJ)) ||
        (rc = dev_alloc(&c->d_ci_count, (size_t)ncl_max + 4)) ||
        (rc = dev_alloc(&c->d_ci_start, (size_t)ncl_max + 4)) ||
        (rc = dev_alloc(&c->d_wcount, (size_t)ncl_max + 4)) ||
        (rc = dev_alloc(&c->d_wstart, (size_t)ncl_max + 4)) ||
        (rc = dev_alloc(&c->d_witem, (size_t)ncl_max * (PGK_MAXJ / 64))) ||
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
    c->ncl_cap = ncl_max;
    c->acc_dirty = true;   // new accumulators start zeroed by the next call
  }
  if (c->cap_pair_cl == 0) {
    // first guess; grown below when the search reports more entries
    const int64_t cap = (int64_t)ncl_max * 128 + 1024;
    if ((rc = dev_alloc(&c->d_cj, cap)) || (rc = dev_alloc(&c->d_mask, cap))) return rc;
    c->cap_pair_cl = cap;
  }
  if (!c->d_list_meta) {
    HIP_CHECK(hipMalloc(&c->d_list_meta, 4 * sizeof(int32_t)));
    HIP_CHECK(hipHostMalloc(&c->h_list_meta, 4 * sizeof(int32_t)));
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
  const int32_t *d_ncl = c->d_col_clstart + ncolumn;
  k_fill_clusters<<<ncolumn, 64, 0, st>>>(g.nzb, c->d_bin_start, c->d_sorted, c->d_col_ncl,
                                          c->d_col_clstart, c->d_cl_atom, c->d_cl_col,
                                          c->d_slot_of);
  k_bbox<<<pgk_grid_blocks(ncl_max, 128), 128, 0, st>>>(d_ncl, c->d_pos, c->d_cl_atom, g,
                                                        c->d_bbox, c->d_posc);
  HIP_CHECK(hipGetLastError());

  ListArgs a;
  a.posc = c->d_posc;
  a.xs_start = c->d_xl_start;
  a.xs_j = c->d_xl_j;
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
  a.ncolumn = ncolumn;
  a.ncl_max = ncl_max;
  for (int d = 0; d < 2; d++) {
    a.ncol[d] = g.ncol[d];
    a.rng[d] = (int)ceil(rcut_book / g.colw[d]);
    a.colwf[d] = (float)g.colw[d];
  }
  for (int d = 0; d < 3; d++) {
    a.L[d] = g.L[d];
    a.Linv[d] = 1.0 / g.L[d];
    a.Lf[d] = (float)g.L[d];
  }
  a.sqcut = rcut_book * rcut_book;
  float rf = (float)rcut_book + 1e-3f;
  a.rcbf2 = rf * rf;
  // reach plus one z bin (cluster z bounds are ordered only up to a bin) and
  // a margin for the bounding-box padding and float rounding
  a.zwin = rf + (float)g.zbw + 0.01f;
  // blocks cover ci = 0 .. ncl_max so that the count past the end is cleared
  if ((int64_t)ncl_max < c->pair_chunk_below)
    k_find_pairs<1><<<ncl_max + 1, 64, 0, st>>>(a);
  else
    k_find_pairs<4><<<(ncl_max + 1 + 3) / 4, 256, 0, st>>>(a);
  HIP_CHECK(hipGetLastError());
  tb = c->scan_tmp_bytes;
  HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(c->d_scan_tmp, tb, c->d_ci_count, c->d_ci_start,
                                             ncl_max + 1, st));
  k_compact_pairs<<<(ncl_max + 3) / 4, 256, 0, st>>>(ncl_max, c->d_ci_count, c->d_ci_start,
                                                     c->d_tmp_cj, c->d_tmp_mask, c->d_cj,
                                                     c->d_mask, c->cap_pair_cl, c->d_overflow);
  k_chunk_count<<<pgk_grid_blocks(ncl_max + 1, 256), 256, 0, st>>>(ncl_max, c->d_ci_count,
                                                                   c->d_wcount);
  tb = c->scan_tmp_bytes;
  HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(c->d_scan_tmp, tb, c->d_wcount, c->d_wstart,
                                             ncl_max + 1, st));
  k_chunk_fill<<<pgk_grid_blocks(ncl_max, 256), 256, 0, st>>>(
      ncl_max, c->d_wcount, c->d_wstart, c->d_witem, (int64_t)ncl_max * (PGK_MAXJ / 64));
  k_list_meta<<<1, 64, 0, st>>>(d_ncl, c->d_overflow, c->d_ci_start + ncl_max,
                                c->d_wstart + ncl_max, c->d_list_meta);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(c->h_list_meta, c->d_list_meta, 4 * sizeof(int32_t),
                           hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  const int32_t ncl = c->h_list_meta[0], ovf = c->h_list_meta[1], tot = c->h_list_meta[2];
  c->nwi = c->h_list_meta[3];
  if (ovf & 1) {
    pgk_set_error("pgk_build_list: more than 2048 neighbour candidates or half-cluster "
                  "entries for one cluster (rcut_book too large for the cluster list)");
    return PGK_EOVERFLOW;
  }
  if (ovf & 2) {
    const int64_t cap = (int64_t)(tot * 1.25) + 1024;
    if ((rc = dev_alloc(&c->d_cj, cap)) || (rc = dev_alloc(&c->d_mask, cap))) return rc;
    c->cap_pair_cl = cap;
    k_compact_pairs<<<(ncl_max + 3) / 4, 256, 0, st>>>(ncl_max, c->d_ci_count, c->d_ci_start,
                                                       c->d_tmp_cj, c->d_tmp_mask, c->d_cj,
                                                       c->d_mask, c->cap_pair_cl, c->d_overflow);
    HIP_CHECK(hipGetLastError());
  }
  c->ncl = ncl;
  c->ncol[0] = g.ncol[0];
  c->ncol[1] = g.ncol[1];
  c->nzb = g.nzb;
  c->npair_cl = tot;
  c->npairs = -1;   // counted on request
  c->have_list = true;
  c->list_serial++;
  for (int d = 0; d < 3; d++) c->list_box[d] = c->box[d];
  return PGK_OK;
}

int pgk_list_count_pairs(pgk_ctx *c) {
  if (c->npairs >= 0) return PGK_OK;
  hipStream_t st = c->stream;
  unsigned long long *d_np = reinterpret_cast<unsigned long long *>(c->d_list_meta);
  HIP_CHECK(hipMemsetAsync(d_np, 0, sizeof(unsigned long long), st));
  if (c->npair_cl > 0) {
    const int nb = (int)std::min<int64_t>(1024, (c->npair_cl + 255) / 256);
    k_count_pairs<<<nb, 256, 0, st>>>(c->npair_cl, c->d_mask, d_np);
    HIP_CHECK(hipGetLastError());
  }
  unsigned long long np = 0;
  HIP_CHECK(hipMemcpyAsync(&np, d_np, sizeof(np), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  c->npairs = (int64_t)np;
  return PGK_OK;
}
