/*
 * pgk_oracle.c - plain-C restatement of peachgk_md's force-evaluation path.
 *
 * TEST INFRASTRUCTURE: this is the checker the GPU library is compared with.
 * It is never linked into the product.  Each routine restates the reference
 * routine named in its comment, in the same floating-point evaluation order
 * (FP64, no fused multiply-add: build with -ffp-contract=off), so that its
 * results agree with the compiled reference to rounding.
 *
 * Conventions: arrays hold Fortran layouts (atmcor(3,natom) as pos[3*i+d]),
 * atom indices stored 1-based as in the reference, 0 marks an empty CSR run.
 *
 * Pinning: every function here is compared with the reference compiled from
 * source (oracle/_ref) and with closed forms in tests/test_oracle_*.py.
 */
#include "pgk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXORD 10

struct pgko_sys {
  int32_t natom, ntype;
  double *chrg, *eps, *rmin;
  int32_t *atype, *molid, *inttyp;
  int32_t nexcl;
  int32_t *excl_index, *excl_list;
  int32_t nbond, nbondtyp, *bond_ij, *bond_type;
  double *bond_k, *bond_r0;
  int32_t nangl, nangltyp, *angl_ijk, *angl_type;
  double *angl_k, *angl_t0;
  int32_t ntors, ntorstyp, *tors_ijkl, *tors_type, *tors_divfac, *tors_if14;
  double *tors_ring, *tors_barhig, *tors_phsang, *tors_period, *tors_dvdw,
      *tors_delc;
  int32_t ntorsrb, ntorsrbtyp, *rb_ijkl, *rb_type, *rb_if14;
  double *rb_ring, *rb_c, *rb_dvdw, *rb_delc;
  double alpha, rrcut, rcut;
  int32_t nspl, nfft[3], order, ifnetq;
  /* state */
  double box[3];
  double *pos;
  double netchrgsq, pot_ewc, rkcut;
  /* spline tables: x, and (a,b,c,d) for real, realpot, excs, excspot */
  double spl_int;
  double *x_tbl;
  double *tb[16];
  double *bsp_mod[3];
  /* neighbour list */
  int64_t nlist;
  int32_t *nb_index; /* [natom+1] */
  int32_t *nb_list;
};

/* ------------------------------------------------------------------------ */
static void *dupa(const void *p, size_t n) {
  if (!p || n == 0) return calloc(1, n ? n : 8);
  void *q = malloc(n);
  memcpy(q, p, n);
  return q;
}

/* Fortran dnint / anint: round half away from zero */
static inline double fnint(double x) { return round(x); }

/* ---- spline tables: spline_interp.F90:14-190 --------------------------- */
static void cal_spl_coeff(const pgko_sys *s, const double *y, double *b,
                          double *c, double *d) {
  int n = s->nspl;
  double h = s->spl_int;
  double *al = calloc(n + 1, sizeof(double)), *l = calloc(n + 1, sizeof(double));
  double *mu = calloc(n + 1, sizeof(double)), *z = calloc(n + 1, sizeof(double));
  /* 1-based loops as spline_interp.F90:154-188 */
  for (int i = 2; i <= n - 1; i++)
    al[i] = 3.0 / h * ((y[i] - y[i - 1]) - (y[i - 1] - y[i - 2]));
  l[1] = 1.0;
  mu[1] = 0.0;
  z[1] = 0.0;
  for (int i = 2; i <= n - 1; i++) {
    l[i] = 2.0 * (s->x_tbl[i] - s->x_tbl[i - 2]) - h * mu[i - 1];
    mu[i] = h / l[i];
    z[i] = (al[i] - h * z[i - 1]) / l[i];
  }
  l[n] = 1.0;
  z[n] = 0.0;
  c[n - 1] = 0.0;
  for (int i = n - 1; i >= 1; i--) {
    c[i - 1] = z[i] - mu[i] * c[i];
    b[i - 1] = (y[i] - y[i - 1]) / h - h * (c[i] + 2.0 * c[i - 1]) / 3.0;
    d[i - 1] = (c[i] - c[i - 1]) / (3.0 * h);
  }
  free(al); free(l); free(mu); free(z);
}

static void spline_init(pgko_sys *s) {
  int n = s->nspl;
  double root_pi = sqrt(acos(-1.0));
  double pie = 2.0 / root_pi * s->alpha;
  s->spl_int = (s->rrcut + 2.0) / (double)n;
  s->x_tbl = calloc(n, sizeof(double));
  for (int k = 0; k < 16; k++) s->tb[k] = calloc(n, sizeof(double));
  for (int i = 0; i < n; i++) s->x_tbl[i] = (double)i * s->spl_int;
  double *re = s->tb[0], *rp = s->tb[4], *ex = s->tb[8], *ep = s->tb[12];
  re[0] = rp[0] = ex[0] = ep[0] = 1.0;
  for (int i = 1; i < n; i++) {
    double r = s->x_tbl[i], inv = 1.0 / r, arg = s->alpha * r;
    re[i] = inv * (pie * exp(-arg * arg) * inv + erfc(arg) * inv * inv);
    rp[i] = erfc(arg) * inv;
    ex[i] = inv * (-pie * exp(-arg * arg) * inv + (1.0 - erfc(arg)) * inv * inv);
    ep[i] = (1.0 - erfc(arg)) * inv;
  }
  for (int t = 0; t < 4; t++)
    cal_spl_coeff(s, s->tb[4 * t], s->tb[4 * t + 1], s->tb[4 * t + 2],
                  s->tb[4 * t + 3]);
}

/* SPLINE_INTERPFUNC: spline_interp.F90:196-242 (t = table set 0..3) */
static inline double spl(const pgko_sys *s, int t, double x) {
  if (x >= s->spl_int * (double)(s->nspl - 1)) return 0.0;
  int i = (int)(x / s->spl_int);
  double d = x - s->x_tbl[i];
  const double *a = s->tb[4 * t], *b = s->tb[4 * t + 1], *c = s->tb[4 * t + 2],
               *dd = s->tb[4 * t + 3];
  return a[i] + d * (b[i] + d * (c[i] + d * dd[i]));
}

/* ---- SPME set-up: fft_pme_init.F90:216-289, erf_corr_cutoff.F90:28-62 --- */
static void fill_bspline(double w, int order, double *array, double *darray) {
  /* arrays 1-based in the reference; here array[0..order-1] */
  double *A = array - 1, *D = darray - 1;
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

static void dftmod(double *mod, const double *arr, int n) {
  double twopi = 2.0 * 3.14159265358979323846, tiny = 1.0e-7;
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

static void pme_init(pgko_sys *s, const double *box0) {
  double arr[MAXORD], darr[MAXORD];
  int maxn = s->nfft[0];
  if (s->nfft[1] > maxn) maxn = s->nfft[1];
  if (s->nfft[2] > maxn) maxn = s->nfft[2];
  double *bsp = calloc(maxn + MAXORD + 2, sizeof(double));
  fill_bspline(0.0, s->order, arr, darr);
  for (int i = 2; i <= s->order + 1; i++) bsp[i - 1] = arr[i - 2];
  for (int d = 0; d < 3; d++) {
    s->bsp_mod[d] = calloc(s->nfft[d], sizeof(double));
    dftmod(s->bsp_mod[d], bsp, s->nfft[d]);
  }
  free(bsp);
  /* self energy fft_pme_init.F90:149-156 */
  double pi = acos(-1.0), e = 0.0;
  for (int i = 0; i < s->natom; i++) e = e + s->chrg[i] * s->chrg[i];
  s->pot_ewc = -1.0 / sqrt(pi) * s->alpha * e;
  /* rkcut with the reference's comparison order (erf_corr_cutoff.F90:49-57) */
  double m1 = 0.5 * (s->nfft[0] / box0[0]);
  double rk = m1;
  double m2 = 0.5 * (s->nfft[1] / box0[1]);
  if (m2 < m1) rk = m2;
  double m3 = 0.5 * (s->nfft[2] / box0[2]);
  if (m3 < m2) rk = m3;
  s->rkcut = 0.9 * rk;
}

/* ------------------------------------------------------------------------ */
pgko_sys *pgko_create(const pgk_system_desc *d) {
  pgko_sys *s = calloc(1, sizeof(*s));
  int n = d->natom, nt = d->ntype;
  s->natom = n;
  s->ntype = nt;
  s->chrg = dupa(d->charge, sizeof(double) * n);
  s->atype = dupa(d->atype, sizeof(int32_t) * n);
  s->molid = dupa(d->molid, sizeof(int32_t) * n);
  s->eps = dupa(d->lj_eps, sizeof(double) * nt * nt);
  s->rmin = dupa(d->lj_rmin, sizeof(double) * nt * nt);
  s->inttyp = dupa(d->inttyp, sizeof(int32_t) * nt * nt);
  s->nexcl = d->nexcl;
  s->excl_index = dupa(d->excl_index, sizeof(int32_t) * (n + 1));
  s->excl_list = dupa(d->excl_list, sizeof(int32_t) * (d->nexcl > 0 ? d->nexcl : 1));
  s->nbond = d->nbond; s->nbondtyp = d->nbondtyp;
  s->bond_ij = dupa(d->bond_ij, sizeof(int32_t) * 2 * d->nbond);
  s->bond_type = dupa(d->bond_type, sizeof(int32_t) * d->nbond);
  s->bond_k = dupa(d->bond_k, sizeof(double) * d->nbondtyp);
  s->bond_r0 = dupa(d->bond_r0, sizeof(double) * d->nbondtyp);
  s->nangl = d->nangl; s->nangltyp = d->nangltyp;
  s->angl_ijk = dupa(d->angl_ijk, sizeof(int32_t) * 3 * d->nangl);
  s->angl_type = dupa(d->angl_type, sizeof(int32_t) * d->nangl);
  s->angl_k = dupa(d->angl_k, sizeof(double) * d->nangltyp);
  s->angl_t0 = dupa(d->angl_theta0, sizeof(double) * d->nangltyp);
  s->ntors = d->ntors; s->ntorstyp = d->ntorstyp;
  s->tors_ijkl = dupa(d->tors_ijkl, sizeof(int32_t) * 4 * d->ntors);
  s->tors_type = dupa(d->tors_type, sizeof(int32_t) * d->ntors);
  s->tors_ring = dupa(d->tors_ring14, sizeof(double) * d->ntors);
  s->tors_divfac = dupa(d->tors_divfac, sizeof(int32_t) * d->ntorstyp);
  s->tors_barhig = dupa(d->tors_barhig, sizeof(double) * d->ntorstyp);
  s->tors_phsang = dupa(d->tors_phsang, sizeof(double) * d->ntorstyp);
  s->tors_period = dupa(d->tors_period, sizeof(double) * d->ntorstyp);
  s->tors_dvdw = dupa(d->tors_divfac_vdw, sizeof(double) * d->ntorstyp);
  s->tors_delc = dupa(d->tors_divfac_elc, sizeof(double) * d->ntorstyp);
  s->tors_if14 = dupa(d->tors_if14, sizeof(int32_t) * d->ntorstyp);
  s->ntorsrb = d->ntorsrb; s->ntorsrbtyp = d->ntorsrbtyp;
  s->rb_ijkl = dupa(d->torsrb_ijkl, sizeof(int32_t) * 4 * d->ntorsrb);
  s->rb_type = dupa(d->torsrb_type, sizeof(int32_t) * d->ntorsrb);
  s->rb_ring = dupa(d->torsrb_ring14, sizeof(double) * d->ntorsrb);
  s->rb_c = dupa(d->torsrb_c, sizeof(double) * 6 * d->ntorsrbtyp);
  s->rb_dvdw = dupa(d->torsrb_divfac_vdw, sizeof(double) * d->ntorsrbtyp);
  s->rb_delc = dupa(d->torsrb_divfac_elc, sizeof(double) * d->ntorsrbtyp);
  s->rb_if14 = dupa(d->torsrb_if14, sizeof(int32_t) * d->ntorsrbtyp);
  s->alpha = d->alpha;
  s->rrcut = d->rrcut;
  s->rcut = d->rcut;
  s->nspl = d->nspltbl;
  s->nfft[0] = d->nfft1; s->nfft[1] = d->nfft2; s->nfft[2] = d->nfft3;
  s->order = d->pme_order;
  s->ifnetq = d->ifnetqcorrp;
  memcpy(s->box, d->box, sizeof(s->box));
  s->pos = calloc(3 * (size_t)n + 3, sizeof(double));
  /* netchrgsq: prepmd.F90:244-249 */
  double q = 0.0;
  for (int i = 0; i < n; i++) q = q + s->chrg[i];
  s->netchrgsq = q * q;
  spline_init(s);
  pme_init(s, d->box);
  return s;
}

void pgko_destroy(pgko_sys *s) {
  if (!s) return;
  void *ptrs[] = {s->chrg, s->eps, s->rmin, s->atype, s->molid, s->inttyp,
                  s->excl_index, s->excl_list, s->bond_ij, s->bond_type,
                  s->bond_k, s->bond_r0, s->angl_ijk, s->angl_type, s->angl_k,
                  s->angl_t0, s->tors_ijkl, s->tors_type, s->tors_ring,
                  s->tors_divfac, s->tors_barhig, s->tors_phsang,
                  s->tors_period, s->tors_dvdw, s->tors_delc, s->tors_if14,
                  s->rb_ijkl, s->rb_type, s->rb_ring, s->rb_c, s->rb_dvdw,
                  s->rb_delc, s->rb_if14, s->pos, s->x_tbl, s->nb_index,
                  s->nb_list, s->bsp_mod[0], s->bsp_mod[1], s->bsp_mod[2]};
  for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); k++) free(ptrs[k]);
  for (int k = 0; k < 16; k++) free(s->tb[k]);
  free(s);
}

void pgko_set_box(pgko_sys *s, const double *box) { memcpy(s->box, box, 3 * sizeof(double)); }
void pgko_set_positions(pgko_sys *s, const double *pos) {
  memcpy(s->pos, pos, 3 * sizeof(double) * (size_t)s->natom);
}
double pgko_pot_ewc(const pgko_sys *s) { return s->pot_ewc; }

void pgko_get_tables(const pgko_sys *s, double *tbl, double *bsp, double *rkcut) {
  for (int k = 0; k < 16; k++) memcpy(tbl + (size_t)k * s->nspl, s->tb[k], sizeof(double) * s->nspl);
  int o = 0;
  for (int d = 0; d < 3; d++) {
    memcpy(bsp + o, s->bsp_mod[d], sizeof(double) * s->nfft[d]);
    o += s->nfft[d];
  }
  *rkcut = s->rkcut;
}

/* ---- exclusions: mkexcl.F90:6-223 --------------------------------------- */
static int cmp_i32(const void *a, const void *b) {
  int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

int32_t pgko_mkexcl(const pgk_system_desc *d, int32_t *index_out,
                    int32_t *list_out, int32_t capacity) {
  int n = d->natom;
  /* count partners per atom (a superset; duplicates dropped after sort) */
  int32_t *cnt = calloc(n + 2, sizeof(int32_t));
#define ADDP(a, b) do { cnt[(a)]++; cnt[(b)]++; } while (0)
  for (int k = 0; k < d->nbond; k++) ADDP(d->bond_ij[2 * k], d->bond_ij[2 * k + 1]);
  for (int k = 0; k < d->nangl; k++) ADDP(d->angl_ijk[3 * k], d->angl_ijk[3 * k + 2]);
  for (int k = 0; k < d->ntors; k++) ADDP(d->tors_ijkl[4 * k], d->tors_ijkl[4 * k + 3]);
  for (int k = 0; k < d->ntorsrb; k++) ADDP(d->torsrb_ijkl[4 * k], d->torsrb_ijkl[4 * k + 3]);
#undef ADDP
  int64_t *start = calloc(n + 2, sizeof(int64_t));
  for (int i = 1; i <= n; i++) start[i + 1] = start[i] + cnt[i];
  int32_t *buf = calloc(start[n + 1] + 1, sizeof(int32_t));
  int32_t *fill = calloc(n + 2, sizeof(int32_t));
#define PUT(a, b) do { buf[start[(a)] + fill[(a)]++] = (b); buf[start[(b)] + fill[(b)]++] = (a); } while (0)
  for (int k = 0; k < d->nbond; k++) PUT(d->bond_ij[2 * k], d->bond_ij[2 * k + 1]);
  for (int k = 0; k < d->nangl; k++) PUT(d->angl_ijk[3 * k], d->angl_ijk[3 * k + 2]);
  for (int k = 0; k < d->ntors; k++) PUT(d->tors_ijkl[4 * k], d->tors_ijkl[4 * k + 3]);
  for (int k = 0; k < d->ntorsrb; k++) PUT(d->torsrb_ijkl[4 * k], d->torsrb_ijkl[4 * k + 3]);
#undef PUT
  int32_t nl = 0;
  for (int i = 1; i <= n; i++) {
    int32_t *p = buf + start[i];
    int m = fill[i];
    qsort(p, m, sizeof(int32_t), cmp_i32);
    index_out[i - 1] = nl + 1;
    int any = 0;
    int32_t last = -1;
    for (int k = 0; k < m; k++) {
      if (p[k] > i && p[k] != last) {
        if (nl < capacity) list_out[nl] = p[k];
        nl++;
        any = 1;
      }
      last = p[k];
    }
    if (!any) {
      if (nl < capacity) list_out[nl] = 0;
      nl++;
    }
  }
  index_out[n] = nl + 1;
  free(cnt); free(start); free(buf); free(fill);
  return nl;
}

/* ---- Verlet list: mklist.F90:8-234 and mklist_cell.F90:8-352 ------------ */
static int is_excluded(const pgko_sys *s, int i, int j) { /* i<j, 1-based */
  for (int k = s->excl_index[i - 1]; k < s->excl_index[i]; k++)
    if (s->excl_list[k - 1] == j) return 1;
  return 0;
}

static inline double dist2_ref(const pgko_sys *s, int i, int j) {
  /* rij = x_i - x_j, minimum image with dnint, |rij|^2 without FMA */
  double r[3];
  for (int d = 0; d < 3; d++) {
    double bi = 1.0 / s->box[d];
    r[d] = s->pos[3 * (i - 1) + d] - s->pos[3 * (j - 1) + d];
    r[d] = r[d] - s->box[d] * fnint(r[d] * bi);
  }
  return r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
}

int64_t pgko_mklist(pgko_sys *s, double rcut_book, int32_t mode) {
  int n = s->natom;
  double sqcut = rcut_book * rcut_book;
  size_t cap = 1024 + (size_t)n * 64;
  int32_t *lst = malloc(cap * sizeof(int32_t));
  int32_t *idx = calloc(n + 1, sizeof(int32_t));
  int32_t *tmp = malloc(sizeof(int32_t) * (n + 1));
  int64_t icut = 0;
  int cc[3] = {0, 0, 0};
  int32_t *cell_start = NULL, *cell_atoms = NULL, *a2c = NULL;
  if (mode == 1) {
    /* cell index (mklist_cell.F90:113-176) */
    double csz[3];
    for (int d = 0; d < 3; d++) {
      cc[d] = (int)(s->box[d] / rcut_book);
      if (cc[d] < 1) cc[d] = 1;
      csz[d] = s->box[d] / (double)cc[d];
    }
    int ncell = cc[0] * cc[1] * cc[2];
    /* MOD(INT((x + L)/cellsize), ncell) indexes a cell only for x >= -L:
     * coordinates below that are outside the reference's domain */
    for (int i = 0; i < 3 * n; i++)
      if (!(s->pos[i] + s->box[i % 3] >= 0.0)) {
        free(lst); free(idx); free(tmp);
        return -2;
      }
    a2c = malloc(sizeof(int32_t) * 3 * n);
    int32_t *ccount = calloc(ncell + 1, sizeof(int32_t));
    int32_t *cid = malloc(sizeof(int32_t) * n);
    for (int i = 0; i < n; i++) {
      for (int d = 0; d < 3; d++)
        a2c[3 * i + d] = ((int)((s->pos[3 * i + d] + s->box[d]) / csz[d])) % cc[d];
      cid[i] = a2c[3 * i] + cc[0] * (a2c[3 * i + 1] + cc[1] * a2c[3 * i + 2]);
      ccount[cid[i] + 1]++;
    }
    cell_start = calloc(ncell + 1, sizeof(int32_t));
    for (int c = 0; c < ncell; c++) cell_start[c + 1] = cell_start[c] + ccount[c + 1];
    cell_atoms = malloc(sizeof(int32_t) * (n + 1));
    int32_t *f = calloc((size_t)ncell + 1, sizeof(int32_t));
    for (int i = 0; i < n; i++) cell_atoms[cell_start[cid[i]] + f[cid[i]]++] = i + 1;
    free(f); free(ccount); free(cid);
  }
  for (int i = 1; i <= n - 1; i++) {
    int itmp = 0;
    if (mode == 0) {
      for (int j = i + 1; j <= n; j++)
        if (dist2_ref(s, i, j) < sqcut) tmp[itmp++] = j;
    } else {
      const int32_t *ce = a2c + 3 * (i - 1);
      for (int z1 = ce[2] - 1; z1 <= ce[2] + 1; z1++) {
        int z = (z1 + cc[2]) % cc[2];
        for (int y1 = ce[1] - 1; y1 <= ce[1] + 1; y1++) {
          int y = (y1 + cc[1]) % cc[1];
          int posyz = y * cc[0] + z * cc[0] * cc[1];
          for (int x1 = ce[0] - 1; x1 <= ce[0] + 1; x1++) {
            int x = (x1 + cc[0]) % cc[0];
            int c = x + posyz;
            for (int k = cell_start[c]; k < cell_start[c + 1]; k++) {
              int j = cell_atoms[k];
              if (i >= j) continue;
              if (dist2_ref(s, i, j) < sqcut) tmp[itmp++] = j;
            }
          }
        }
      }
    }
    int iitmp = itmp;
    for (int k = 0; k < itmp; k++)
      if (is_excluded(s, i, tmp[k])) { tmp[k] = -999; iitmp--; }
    if ((size_t)(icut + itmp + 1) > cap) {
      cap = 2 * cap + itmp + 1;
      lst = realloc(lst, cap * sizeof(int32_t));
    }
    idx[i - 1] = (int32_t)(icut + 1);
    if (iitmp == 0) {
      lst[icut++] = 0;
    } else {
      for (int k = 0; k < itmp; k++)
        if (tmp[k] > 0) lst[icut++] = tmp[k];
    }
  }
  idx[n - 1] = (int32_t)(icut + 1); /* nb_indexall(natom) (mklist.F90:191) */
  idx[n] = (int32_t)(icut + 1);
  free(tmp); free(a2c); free(cell_start); free(cell_atoms);
  free(s->nb_index); free(s->nb_list);
  s->nb_index = idx;
  s->nb_list = lst;
  s->nlist = icut;
  return icut;
}

void pgko_get_list(const pgko_sys *s, int32_t *index_out, int32_t *list_out) {
  memcpy(index_out, s->nb_index, sizeof(int32_t) * s->natom);
  memcpy(list_out, s->nb_list, sizeof(int32_t) * s->nlist);
}

void pgko_set_list(pgko_sys *s, int64_t nl, const int32_t *index_in,
                   const int32_t *list_in) {
  free(s->nb_index); free(s->nb_list);
  s->nb_index = malloc(sizeof(int32_t) * (s->natom + 1));
  memcpy(s->nb_index, index_in, sizeof(int32_t) * s->natom);
  s->nb_index[s->natom] = (int32_t)(nl + 1);
  s->nb_list = malloc(sizeof(int32_t) * (nl > 0 ? nl : 1));
  memcpy(s->nb_list, list_in, sizeof(int32_t) * nl);
  s->nlist = nl;
}

/* ---- heat-flux virial scatter: calvirihf.F90:8-405 ---------------------- */
typedef struct hfctx {
  const pgk_hf_desc *hf;
  int natom;
  double *w[PGK_NHF];  /* (3,3,natom,nreg) per class */
  double *pe[PGK_NHF]; /* (natom) per class */
} hfctx;

#define WIX(a, n, i, r, N) ((a) + 3 * (n) + 9 * ((size_t)(i) + (size_t)(N) * (r)))

static void calvirihf(const pgko_sys *s, const hfctx *h, int i, int j,
                      const double *fi, const double *fj, double nbc,
                      int cls) {
  const pgk_hf_desc *hf = h->hf;
  const double *box = s->box;
  double binv[3] = {1.0 / box[0], 1.0 / box[1], 1.0 / box[2]};
  double ci[3], cj[3];
  int atomflag = 0;
  int bulk = (hf->mode == PGK_HF_BULK);
  if (hf->hftyp[i - 1] == PGK_HFTYP_MOLE) {
    int m = s->molid[i - 1];
    for (int d = 0; d < 3; d++) ci[d] = hf->molecom[3 * (m - 1) + d];
  } else {
    atomflag++;
    for (int d = 0; d < 3; d++) ci[d] = s->pos[3 * (i - 1) + d];
    if (!bulk) {
      if (ci[2] < 0.0) ci[2] = ci[2] + box[2];
      else if (ci[2] >= box[2]) ci[2] = ci[2] - box[2];
    }
  }
  if (hf->hftyp[j - 1] == PGK_HFTYP_MOLE) {
    int m = s->molid[j - 1];
    for (int d = 0; d < 3; d++) cj[d] = hf->molecom[3 * (m - 1) + d];
  } else {
    atomflag++;
    for (int d = 0; d < 3; d++) cj[d] = s->pos[3 * (j - 1) + d];
    if (!bulk) {
      if (cj[2] < 0.0) cj[2] = cj[2] + box[2];
      else if (cj[2] >= box[2]) cj[2] = cj[2] - box[2];
    }
  }
  if (atomflag == 0 && s->molid[i - 1] == s->molid[j - 1]) return;
  double r[3] = {0.0, 0.0, 0.0}, rz = 0.0;
  if (bulk) {
    r[0] = ci[0] - cj[0];
    r[1] = ci[1] - cj[1];
  }
  r[2] = ci[2] - cj[2];
  if (!bulk) {
    rz = fabs(r[2]);
    if (rz * binv[2] >= 0.5) return;
  }
  double *W = h->w[cls];
  int N = h->natom;
  if (hf->mode == PGK_HF_SURFACE) {
    for (int ir = 0; ir < hf->nregion; ir++) {
      double ri1 = ci[2] - hf->z1[ir], rj1 = cj[2] - hf->z1[ir];
      if (ri1 * rj1 < 0.0) {
        double ff = (ri1 > 0.0) ? 1.0 : -1.0;
        ff = ff * nbc;
        for (int a = 0; a < 3; a++) {
          WIX(W, 2, i - 1, ir, N)[a] += fi[a] * ff;
          WIX(W, 2, j - 1, ir, N)[a] += (-fj[a]) * ff;
        }
      }
    }
    return;
  }
  if (bulk) {
    for (int d = 0; d < 3; d++) r[d] = r[d] - box[d] * fnint(r[d] * binv[d]);
    for (int ir = 0; ir < hf->nregion; ir++) {
      double ff = nbc;
      for (int n = 0; n < 3; n++)
        for (int a = 0; a < 3; a++) {
          double rfi = r[n] * fi[a], rfj = r[n] * (-fj[a]);
          WIX(W, n, i - 1, ir, N)[a] += rfi * ff;
          WIX(W, n, j - 1, ir, N)[a] += rfj * ff;
        }
    }
    return;
  }
  /* default volume mode */
  for (int ir = 0; ir < hf->nregion; ir++) {
    double z1 = hf->z1[ir], z2 = hf->z2[ir];
    double ri1 = ci[2] - z1, rj1 = cj[2] - z1, ri2 = ci[2] - z2, rj2 = cj[2] - z2;
    int np = 0, p1 = 0;
    if (ri1 * rj1 < 0.0) { p1 = 1; np++; }
    if (ri2 * rj2 < 0.0) { np++; }
    double ff;
    if (np == 0) {
      if (ri1 * ri2 > 0.0) continue;
      ff = rz;
    } else if (np == 1) {
      if (p1) ff = (ri1 > 0.0) ? ri1 : rj1;
      else ff = (ri2 < 0.0) ? -ri2 : -rj2;
    } else {
      ff = z2 - z1;
    }
    if (r[2] >= 0.0) ff = ff * nbc;
    else ff = -ff * nbc;
    for (int a = 0; a < 3; a++) {
      WIX(W, 2, i - 1, ir, N)[a] += fi[a] * ff;
      WIX(W, 2, j - 1, ir, N)[a] += (-fj[a]) * ff;
    }
  }
}

/* ---- outputs ------------------------------------------------------------ */
typedef struct acc {
  double *force;          /* [3N] total */
  double *fc, *fl;        /* for_viri_coul / for_viri_lj, may be NULL */
  double ene[PGK_NENE];
  double vt[PGK_NTENS][9];
} acc;

static inline void add3(double *f, int i, const double *v, double sgn) {
  for (int d = 0; d < 3; d++) f[3 * (i - 1) + d] += sgn * v[d];
}
static inline void tens_add(double *t, const double *f, const double *r) {
  /* t(a,n) += f(a)*r(n), Fortran order */
  for (int n = 0; n < 3; n++)
    for (int a = 0; a < 3; a++) t[a + 3 * n] += f[a] * r[n];
}

/* ---- Ewald real space + LJ: ewaldr.F90:404-838, ewaldr_hf.F90:8-670 ---- */
static void ewaldrp(const pgko_sys *s, acc *A, int want_vir, const hfctx *h) {
  int n = s->natom;
  double pi = acos(-1.0);
  double box[3] = {s->box[0], s->box[1], s->box[2]};
  double binv[3] = {1.0 / box[0], 1.0 / box[1], 1.0 / box[2]};
  double sqcut = s->rrcut * s->rrcut, sqcut2 = s->rcut * s->rcut;
  (void)pi;
  if (sqcut2 > sqcut) sqcut = sqcut2;
  double *felc = calloc(3 * (size_t)n, sizeof(double));
  double *fvdw = calloc(3 * (size_t)n, sizeof(double));
  double pot_elc = 0.0, pot_vdw = 0.0;
  double *vtc = A->vt[PGK_T_COUL], *vtl = A->vt[PGK_T_LJ];
  for (int i = 1; i <= n - 1; i++) {
    int j1 = s->nb_index[i - 1], j2 = s->nb_index[i] - 1;
    int ityp = s->atype[i - 1];
    int im = s->molid[i - 1];
    if (s->nb_list[j1 - 1] == 0) continue;
    for (int jx = j1; jx <= j2; jx++) {
      int jj = s->nb_list[jx - 1];
      int jm = s->molid[jj - 1];
      double r[3];
      for (int d = 0; d < 3; d++) {
        r[d] = s->pos[3 * (i - 1) + d] - s->pos[3 * (jj - 1) + d];
        r[d] = r[d] - box[d] * fnint(r[d] * binv[d]);
      }
      double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      if (r2 > sqcut) continue;
      double rinv2 = 1.0 / r2;
      double ra = sqrt(r2);
      double rinv1 = 1.0 / ra;
      if (ra <= s->rrcut) {
        double qq = s->chrg[i - 1] * s->chrg[jj - 1];
        double ft = spl(s, 0, ra) * qq;
        double pot = spl(s, 1, ra) * qq;
        pot_elc = pot_elc + pot;
        double fij[3] = {ft * r[0], ft * r[1], ft * r[2]};
        add3(felc, i, fij, 1.0);
        add3(felc, jj, fij, -1.0);
        if (want_vir) {
          if (A->fc) { add3(A->fc, i, fij, 1.0); add3(A->fc, jj, fij, -1.0); }
          tens_add(vtc, fij, r);
          A->ene[PGK_V_COUL] = A->ene[PGK_V_COUL] + pot;
        }
        if (h) {
          /* default HF Coulomb: bare qq/r (ewaldr_hf.F90:336-343) */
          double ph = qq * rinv1;
          double fh = ph * rinv2;
          double fi[3] = {fh * r[0], fh * r[1], fh * r[2]};
          double fj[3] = {-fi[0], -fi[1], -fi[2]};
          int cls = (im == jm) ? PGK_HF_ELIN : PGK_HF_ELC;
          h->pe[cls][i - 1] += 0.5 * ph;
          h->pe[cls][jj - 1] += 0.5 * ph;
          calvirihf(s, h, i, jj, fi, fj, 0.5, cls);
        }
      }
      if (r2 <= sqcut2) {
        int jtyp = s->atype[jj - 1];
        size_t tij = (size_t)(ityp - 1) + (size_t)s->ntype * (jtyp - 1);
        if (s->inttyp[tij] != 1) continue;
        double rad = s->rmin[tij];
        double t2 = rinv2 * (rad * rad);
        double t6 = t2 * t2 * t2;
        double t12 = t6 * t6;
        double e = s->eps[tij];
        double pot = e * (t12 - 2.0 * t6);
        pot_vdw = pot_vdw + pot;
        double ft = 12.0 * e * (t12 - t6) * rinv2;
        double fij[3] = {ft * r[0], ft * r[1], ft * r[2]};
        add3(fvdw, i, fij, 1.0);
        add3(fvdw, jj, fij, -1.0);
        if (want_vir) {
          if (A->fl) { add3(A->fl, i, fij, 1.0); add3(A->fl, jj, fij, -1.0); }
          tens_add(vtl, fij, r);
          A->ene[PGK_V_LJ] = A->ene[PGK_V_LJ] + fij[0] * r[0] + fij[1] * r[1] + fij[2] * r[2];
        }
        if (h) {
          double fj[3] = {-fij[0], -fij[1], -fij[2]};
          int cls = (im == jm) ? PGK_HF_LJIN : PGK_HF_VDW;
          h->pe[cls][i - 1] += 0.5 * pot;
          h->pe[cls][jj - 1] += 0.5 * pot;
          calvirihf(s, h, i, jj, fij, fj, 0.5, cls);
        }
      }
    }
  }
  /* excluded pairs: subtract qq*erf(alpha r)/r (ewaldr.F90:744-828) */
  for (int i = 1; i <= n - 1; i++) {
    int j1 = s->excl_index[i - 1], j2 = s->excl_index[i] - 1;
    if (s->excl_list[j1 - 1] == 0) continue;
    for (int jx = j1; jx <= j2; jx++) {
      int jj = s->excl_list[jx - 1];
      double r[3];
      for (int d = 0; d < 3; d++) {
        r[d] = s->pos[3 * (i - 1) + d] - s->pos[3 * (jj - 1) + d];
        r[d] = r[d] - box[d] * fnint(r[d] * binv[d]);
      }
      double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      double ra = sqrt(r2);
      double qq = s->chrg[i - 1] * s->chrg[jj - 1];
      double ft = -spl(s, 2, ra) * qq;
      double pot = spl(s, 3, ra) * qq;
      pot_elc = pot_elc - pot;
      double fij[3] = {ft * r[0], ft * r[1], ft * r[2]};
      add3(felc, i, fij, 1.0);
      add3(felc, jj, fij, -1.0);
      if (want_vir) {
        A->ene[PGK_V_COUL] = A->ene[PGK_V_COUL] - pot;
        if (A->fc) { add3(A->fc, i, fij, 1.0); add3(A->fc, jj, fij, -1.0); }
        tens_add(vtc, fij, r);
      }
    }
  }
  for (size_t k = 0; k < 3 * (size_t)n; k++) A->force[k] = A->force[k] + (fvdw[k] + felc[k]);
  A->ene[PGK_E_ELC] = pot_elc;
  A->ene[PGK_E_VDW] = pot_vdw;
  free(felc);
  free(fvdw);
}

/* ---- complex FFT (mixed radix, recursive; stands in for FFTPACK cfftf/b) -- */
static void fft1d(double *re, double *im, int n, int stride, int isign,
                  double *wr, double *wi) {
  /* out-of-place recursive decimation in time into (wr,wi), copied back */
  if (n == 1) return;
  int p = 2;
  while (n % p) p++;
  int m = n / p;
  /* gather sub-sequences x[k*p + r] for r = 0..p-1 */
  double *tr = wr, *ti = wi;
  for (int r = 0; r < p; r++)
    for (int k = 0; k < m; k++) {
      tr[r * m + k] = re[(k * p + r) * stride];
      ti[r * m + k] = im[(k * p + r) * stride];
    }
  for (int r = 0; r < p; r++)
    fft1d(tr + r * m, ti + r * m, m, 1, isign, wr + n, wi + n);
  double twopi = 2.0 * 3.14159265358979323846;
  for (int k = 0; k < m; k++)
    for (int q = 0; q < p; q++) {
      int kk = k + q * m;
      double sr = 0.0, si = 0.0;
      for (int r = 0; r < p; r++) {
        double ang = isign * twopi * (double)((long)r * kk % n) / (double)n;
        double c = cos(ang), sn = sin(ang);
        double xr = tr[r * m + k], xi = ti[r * m + k];
        sr += xr * c - xi * sn;
        si += xr * sn + xi * c;
      }
      re[kk * stride] = sr;
      im[kk * stride] = si;
    }
}

void pgko_fft3d(double *g, int32_t n1, int32_t n2, int32_t n3, int32_t isign) {
  int nmax = n1 > n2 ? n1 : n2;
  if (n3 > nmax) nmax = n3;
  double *lr = malloc(sizeof(double) * nmax), *li = malloc(sizeof(double) * nmax);
  double *wr = malloc(sizeof(double) * 4 * nmax + 64), *wi = malloc(sizeof(double) * 4 * nmax + 64);
  int dims[3] = {n1, n2, n3};
  size_t str[3] = {1, (size_t)n1, (size_t)n1 * n2};
  for (int ax = 0; ax < 3; ax++) {
    int n = dims[ax];
    int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
    for (int b = 0; b < dims[o2]; b++)
      for (int a = 0; a < dims[o1]; a++) {
        size_t base = a * str[o1] + b * str[o2];
        for (int k = 0; k < n; k++) {
          lr[k] = g[2 * (base + k * str[ax])];
          li[k] = g[2 * (base + k * str[ax]) + 1];
        }
        fft1d(lr, li, n, 1, isign, wr, wi);
        for (int k = 0; k < n; k++) {
          g[2 * (base + k * str[ax])] = lr[k];
          g[2 * (base + k * str[ax]) + 1] = li[k];
        }
      }
  }
  free(lr); free(li); free(wr); free(wi);
}

/* ---- SPME: fft_pme.F90:15-270 (fft_pmep), 273-422, 544-690, 872-937 ---- */
static void fft_pmep(const pgko_sys *s, acc *A, int want_vir) {
  int n = s->natom, p = s->order;
  int K1 = s->nfft[0], K2 = s->nfft[1], K3 = s->nfft[2];
  double recip[3] = {1.0 / s->box[0], 1.0 / s->box[1], 1.0 / s->box[2]};
  double vol = s->box[0] * s->box[1] * s->box[2];
  double *fr = malloc(sizeof(double) * 3 * n);
  double *th = malloc(sizeof(double) * 3 * p * n), *dth = malloc(sizeof(double) * 3 * p * n);
  size_t G = (size_t)K1 * K2 * K3;
  double *Q = calloc(2 * G, sizeof(double));
  int K[3] = {K1, K2, K3};
  /* get_scaled_fractionals / get_bspline_coeffs (fft_pme.F90:273-356) */
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      double w = s->pos[3 * i + d] * recip[d];
      w = w - floor(w);
      fr[3 * i + d] = K[d] * w;
      double ww = fr[3 * i + d] - trunc(fr[3 * i + d]);
      fill_bspline(ww, p, th + (size_t)(3 * i + d) * p, dth + (size_t)(3 * i + d) * p);
    }
  /* fill_charge_grid (fft_pme.F90:359-422): index int(fr)-p+1 .. int(fr), wrapped */
  for (int i = 0; i < n; i++) {
    const double *t1 = th + (size_t)(3 * i) * p, *t2 = th + (size_t)(3 * i + 1) * p,
                 *t3 = th + (size_t)(3 * i + 2) * p;
    int k0 = (int)fr[3 * i + 2] - p;
    for (int a3 = 0; a3 < p; a3++) {
      k0++;
      int k = k0 < 0 ? k0 + K3 : k0;
      double prod1 = t3[a3] * s->chrg[i];
      int j0 = (int)fr[3 * i + 1] - p;
      for (int a2 = 0; a2 < p; a2++) {
        j0++;
        int j = j0 < 0 ? j0 + K2 : j0;
        double prod = t2[a2] * prod1;
        int i0 = (int)fr[3 * i] - p;
        for (int a1 = 0; a1 < p; a1++) {
          i0++;
          int ii = i0 < 0 ? i0 + K1 : i0;
          Q[2 * (ii + (size_t)K1 * (j + (size_t)K2 * k))] += t1[a1] * prod;
        }
      }
    }
  }
  pgko_fft3d(Q, K1, K2, K3, -1);
  /* scalar_sum (fft_pme.F90:544-690) */
  double pi = 3.14159265358979323846;
  double fac = pi * pi / (s->alpha * s->alpha);
  double rk2 = s->rkcut * s->rkcut;
  int nf[3];
  for (int d = 0; d < 3; d++) {
    nf[d] = K[d] / 2;
    if (2 * nf[d] < K[d]) nf[d]++;
  }
  double energy = 0.0, vsca = 0.0, vt[3][3] = {{0}};
  for (size_t ind = 1; ind < G; ind++) {
    int k3 = (int)(ind / ((size_t)K1 * K2)) + 1;
    size_t jnd = ind - (size_t)(k3 - 1) * K1 * K2;
    int k2 = (int)(jnd / K1) + 1;
    int k1 = (int)(jnd - (size_t)(k2 - 1) * K1) + 1;
    int m1 = k1 - 1; if (k1 > nf[0]) m1 = k1 - 1 - K1;
    int m2 = k2 - 1; if (k2 > nf[1]) m2 = k2 - 1 - K2;
    int m3 = k3 - 1; if (k3 > nf[2]) m3 = k3 - 1 - K3;
    double h1 = recip[0] * m1, h2 = recip[1] * m2, h3 = recip[2] * m3;
    double msq = h1 * h1 + h2 * h2 + h3 * h3;
    double eterm;
    double *q = Q + 2 * (size_t)(ind);
    if (msq > rk2) {
      eterm = 0.0;
    } else {
      double denom = pi * vol * s->bsp_mod[0][k1 - 1] * s->bsp_mod[1][k2 - 1] *
                     s->bsp_mod[2][k3 - 1] * msq;
      double vterm = 2.0 * (fac * msq + 1.0) / msq;
      double struc2 = q[0] * q[0] + q[1] * q[1];
      double wgt;
      if (!s->ifnetq) {
        eterm = exp(-fac * msq) / denom;
        wgt = eterm * struc2;
      } else {
        double e2 = exp(-fac * msq);
        eterm = e2 / denom;
        e2 = e2 * s->netchrgsq / (pi * vol * msq);
        double e3 = e2 * (-1.0 + 2.0 * fac * msq);
        vsca = vsca + e3;
        wgt = eterm * struc2 - e2;
      }
      energy = energy + eterm * struc2;
      vt[0][0] += wgt * (vterm * h1 * h1 - 1.0);
      vt[0][1] += wgt * (vterm * h1 * h2);
      vt[0][2] += wgt * (vterm * h1 * h3);
      vt[1][1] += wgt * (vterm * h2 * h2 - 1.0);
      vt[1][2] += wgt * (vterm * h2 * h3);
      vt[2][2] += wgt * (vterm * h3 * h3 - 1.0);
    }
    q[0] = eterm * q[0];
    q[1] = eterm * q[1];
  }
  A->ene[PGK_E_EWK] = 0.5 * energy;
  vsca = 0.5 * vsca;
  vt[1][0] = vt[0][1];
  vt[2][0] = vt[0][2];
  vt[2][1] = vt[1][2];
  if (want_vir)
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) A->vt[PGK_T_COUL][a + 3 * b] -= 0.5 * vt[a][b];
  pgko_fft3d(Q, K1, K2, K3, 1);
  /* grad_sum (fft_pme.F90:872-937) */
  for (int i = 0; i < n; i++) {
    const double *t1 = th + (size_t)(3 * i) * p, *t2 = th + (size_t)(3 * i + 1) * p,
                 *t3 = th + (size_t)(3 * i + 2) * p;
    const double *d1 = dth + (size_t)(3 * i) * p, *d2 = dth + (size_t)(3 * i + 1) * p,
                 *d3 = dth + (size_t)(3 * i + 2) * p;
    double f[3] = {0, 0, 0};
    int k0 = (int)fr[3 * i + 2] - p;
    for (int a3 = 0; a3 < p; a3++) {
      k0++;
      int k = k0 < 0 ? k0 + K3 : k0;
      int j0 = (int)fr[3 * i + 1] - p;
      for (int a2 = 0; a2 < p; a2++) {
        j0++;
        int j = j0 < 0 ? j0 + K2 : j0;
        int i0 = (int)fr[3 * i] - p;
        for (int a1 = 0; a1 < p; a1++) {
          i0++;
          int ii = i0 < 0 ? i0 + K1 : i0;
          double term = s->chrg[i] * Q[2 * (ii + (size_t)K1 * (j + (size_t)K2 * k))];
          f[0] = f[0] - K1 * term * d1[a1] * t2[a2] * t3[a3];
          f[1] = f[1] - K2 * term * t1[a1] * d2[a2] * t3[a3];
          f[2] = f[2] - K3 * term * t1[a1] * t2[a2] * d3[a3];
        }
      }
    }
    for (int d = 0; d < 3; d++) {
      A->force[3 * i + d] = A->force[3 * i + d] + recip[d] * f[d];
      if (want_vir && A->fc) A->fc[3 * i + d] = A->fc[3 * i + d] + recip[d] * f[d];
    }
  }
  if (want_vir) A->ene[PGK_V_COUL] = A->ene[PGK_V_COUL] + A->ene[PGK_E_EWK] + vsca;
  free(fr); free(th); free(dth); free(Q);
}

/* ---- bonded terms -------------------------------------------------------- */
static inline void minimg(const pgko_sys *s, double *r) {
  for (int d = 0; d < 3; d++) r[d] = r[d] - s->box[d] * fnint(r[d] * (1.0 / s->box[d]));
}
static inline void vsub(double *r, const double *a, const double *b) {
  for (int d = 0; d < 3; d++) r[d] = a[d] - b[d];
}
static inline double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void cross3(const double *x, const double *y, double *o) {
  o[0] = x[1] * y[2] - x[2] * y[1];
  o[1] = x[2] * y[0] - x[0] * y[2];
  o[2] = x[0] * y[1] - x[1] * y[0];
}
#define POS(i) (s->pos + 3 * ((i) - 1))

/* calbondp: calbond.F90:109-219; calbondp_hf: calbond_hf.F90:6-164 */
static void calbond(const pgko_sys *s, acc *A, int vir, const hfctx *h) {
  double pot = 0.0;
  for (int k = 0; k < s->nbond; k++) {
    int t = s->bond_type[k] - 1;
    if (fabs(s->bond_k[t]) < 1.0e-16) continue;
    int i = s->bond_ij[2 * k], j = s->bond_ij[2 * k + 1];
    double r[3];
    vsub(r, POS(i), POS(j));
    minimg(s, r);
    double ra = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    double dp = s->bond_k[t] * ((ra - s->bond_r0[t]) * (ra - s->bond_r0[t]));
    pot = pot + dp;
    double fa = -2.0 * s->bond_k[t] * (ra - s->bond_r0[t]);
    double f[3] = {fa * r[0] / ra, fa * r[1] / ra, fa * r[2] / ra};
    add3(A->force, i, f, 1.0);
    add3(A->force, j, f, -1.0);
    if (h) {
      double fj[3] = {-f[0], -f[1], -f[2]};
      h->pe[PGK_HF_BO][i - 1] += 0.5 * dp;
      h->pe[PGK_HF_BO][j - 1] += 0.5 * dp;
      calvirihf(s, h, i, j, f, fj, 0.5, PGK_HF_BO);
    }
    if (vir) {
      A->ene[PGK_V_BOND] = A->ene[PGK_V_BOND] + f[0] * r[0] + f[1] * r[1] + f[2] * r[2];
      tens_add(A->vt[PGK_T_BOND], f, r);
    }
  }
  A->ene[PGK_E_BOND] = pot;
}

/* calanglp: calangl.F90:180-358; calanglp_hf: calangl_hf.F90:6-255 */
static void calangl(const pgko_sys *s, acc *A, int vir, const hfctx *h) {
  double pot = 0.0;
  for (int k = 0; k < s->nangl; k++) {
    int t = s->angl_type[k] - 1;
    if (fabs(s->angl_k[t]) < 1.0e-32) continue;
    int i = s->angl_ijk[3 * k], j = s->angl_ijk[3 * k + 1], kk = s->angl_ijk[3 * k + 2];
    double rij[3], rkj[3];
    vsub(rij, POS(j), POS(i));
    minimg(s, rij);
    double ra = sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
    vsub(rkj, POS(j), POS(kk));
    minimg(s, rkj);
    double rb = sqrt(rkj[0] * rkj[0] + rkj[1] * rkj[1] + rkj[2] * rkj[2]);
    double ct = (rij[0] * rkj[0] + rij[1] * rkj[1] + rij[2] * rkj[2]) / (ra * rb);
    if (ct > 1.0) ct = 1.0;
    else if (ct < -1.0) ct = -1.0;
    double th = acos(ct), st = sin(th);
    double dth = th - s->angl_t0[t];
    double dp = s->angl_k[t] * (dth * dth);
    pot = pot + dp;
    double fi[3], fj[3], fk[3];
    if (fabs(st) > 1.0e-14) {
      double ft = -2.0 * s->angl_k[t] * (th - s->angl_t0[t]) / st;
      for (int d = 0; d < 3; d++) {
        fi[d] = (rkj[d] / rb - rij[d] * ct / ra) / ra;
        fi[d] = ft * fi[d];
        fk[d] = (rij[d] / ra - rkj[d] * ct / rb) / rb;
        fk[d] = ft * fk[d];
        fj[d] = -fi[d] - fk[d];
      }
    } else {
      for (int d = 0; d < 3; d++) fi[d] = fj[d] = fk[d] = 0.0;
    }
    add3(A->force, i, fi, 1.0);
    add3(A->force, j, fj, 1.0);
    add3(A->force, kk, fk, 1.0);
    if (h) {
      double c = 1.0 / 3.0;
      h->pe[PGK_HF_AN][i - 1] += c * dp;
      h->pe[PGK_HF_AN][j - 1] += c * dp;
      h->pe[PGK_HF_AN][kk - 1] += c * dp;
      calvirihf(s, h, i, j, fi, fj, c, PGK_HF_AN);
      calvirihf(s, h, j, kk, fj, fk, c, PGK_HF_AN);
      calvirihf(s, h, kk, i, fk, fi, c, PGK_HF_AN);
    }
    if (vir) {
      /* absolute coordinates, as the reference (calangl.F90:231-249) */
      A->ene[PGK_V_ANGL] = A->ene[PGK_V_ANGL] + fi[0] * POS(i)[0] + fi[1] * POS(i)[1] + fi[2] * POS(i)[2];
      A->ene[PGK_V_ANGL] = A->ene[PGK_V_ANGL] + fj[0] * POS(j)[0] + fj[1] * POS(j)[1] + fj[2] * POS(j)[2];
      A->ene[PGK_V_ANGL] = A->ene[PGK_V_ANGL] + fk[0] * POS(kk)[0] + fk[1] * POS(kk)[1] + fk[2] * POS(kk)[2];
      for (int l = 0; l < 3; l++)
        for (int a = 0; a < 3; a++) {
          double *t3 = A->vt[PGK_T_ANGL];
          t3[a + 3 * l] = t3[a + 3 * l] + fi[a] * POS(i)[l];
          t3[a + 3 * l] = t3[a + 3 * l] + fj[a] * POS(j)[l];
          t3[a + 3 * l] = t3[a + 3 * l] + fk[a] * POS(kk)[l];
        }
    }
  }
  A->ene[PGK_E_ANGL] = pot;
}

/* torsion force geometry shared by calijkl / calijklrb (caltors.F90:271-460,
 * caltorsrb.F90:260-459): returns cos(phi), phi and the four forces for a
 * unit dV/dcos-type prefactor f_tmp_a applied by the caller. */
typedef struct torsgeo {
  double rij[3], rkj[3], rkl[3], rik[3], rjl[3];
  double cphi, phi, sphi;
  double nijk[3], njkl[3], nijk_abs, njkl_abs;
} torsgeo;

static void tors_geo(const pgko_sys *s, int i, int j, int k, int l, torsgeo *g) {
  vsub(g->rij, POS(j), POS(i));
  vsub(g->rkj, POS(j), POS(k));
  vsub(g->rkl, POS(l), POS(k));
  vsub(g->rik, POS(k), POS(i));
  vsub(g->rjl, POS(l), POS(j));
  minimg(s, g->rij); minimg(s, g->rkj); minimg(s, g->rkl);
  minimg(s, g->rik); minimg(s, g->rjl);
  cross3(g->rij, g->rkj, g->nijk);
  cross3(g->rkj, g->rkl, g->njkl);
  g->nijk_abs = sqrt(dot3(g->nijk, g->nijk));
  g->njkl_abs = sqrt(dot3(g->njkl, g->njkl));
  double c = dot3(g->nijk, g->njkl) / (g->nijk_abs * g->njkl_abs);
  if (c > 1.0) c = 1.0;
  else if (c < -1.0) c = -1.0;
  double t[3];
  cross3(g->nijk, g->njkl, t);
  double phi = acos(c);
  /* phi = -sign(phi, rkj . (nijk x njkl)) */
  phi = -copysign(phi, dot3(g->rkj, t));
  g->cphi = c;
  g->phi = phi;
  g->sphi = sin(phi);
}

static void tors_forces(torsgeo *g, double fa, double *fi, double *fj,
                        double *fk, double *fl) {
  double ni[3], nj[3], fb[3], fc[3], t1[3], t2[3];
  for (int d = 0; d < 3; d++) {
    ni[d] = g->nijk[d] / g->nijk_abs;
    nj[d] = g->njkl[d] / g->njkl_abs;
  }
  for (int d = 0; d < 3; d++) {
    fb[d] = (nj[d] - g->cphi * ni[d]) / g->nijk_abs;
    fc[d] = (ni[d] - g->cphi * nj[d]) / g->njkl_abs;
  }
  cross3(fb, g->rkj, t1);
  for (int d = 0; d < 3; d++) fi[d] = fa * t1[d];
  cross3(fc, g->rkl, t1);
  cross3(fb, g->rik, t2);
  for (int d = 0; d < 3; d++) fj[d] = fa * (-t1[d] + t2[d]);
  cross3(fb, g->rij, t1);
  cross3(fc, g->rjl, t2);
  for (int d = 0; d < 3; d++) fk[d] = fa * (-t1[d] + t2[d]);
  cross3(fc, g->rkj, t1);
  for (int d = 0; d < 3; d++) fl[d] = fa * t1[d];
}

static void tors_finish(const pgko_sys *s, acc *A, int vir, const hfctx *h,
                        int i, int j, int k, int l, const double *fi,
                        const double *fj, const double *fk, const double *fl,
                        double dp) {
  add3(A->force, i, fi, 1.0);
  add3(A->force, j, fj, 1.0);
  add3(A->force, k, fk, 1.0);
  add3(A->force, l, fl, 1.0);
  if (h) {
    /* caltorsrb_hf.F90 / caltors_hf.F90: 1/4 per atom, six pair scatters */
    double c = 0.25;
    h->pe[PGK_HF_TO][i - 1] += c * dp;
    h->pe[PGK_HF_TO][j - 1] += c * dp;
    h->pe[PGK_HF_TO][k - 1] += c * dp;
    h->pe[PGK_HF_TO][l - 1] += c * dp;
    calvirihf(s, h, i, j, fi, fj, c, PGK_HF_TO);
    calvirihf(s, h, i, k, fi, fk, c, PGK_HF_TO);
    calvirihf(s, h, i, l, fi, fl, c, PGK_HF_TO);
    calvirihf(s, h, j, k, fj, fk, c, PGK_HF_TO);
    calvirihf(s, h, j, l, fj, fl, c, PGK_HF_TO);
    calvirihf(s, h, k, l, fk, fl, c, PGK_HF_TO);
  }
  if (vir) {
    const double *P[4] = {POS(i), POS(j), POS(k), POS(l)};
    const double *F[4] = {fi, fj, fk, fl};
    for (int q = 0; q < 4; q++)
      A->ene[PGK_V_TORS] = A->ene[PGK_V_TORS] + F[q][0] * P[q][0] + F[q][1] * P[q][1] + F[q][2] * P[q][2];
    for (int ll = 0; ll < 3; ll++)
      for (int a = 0; a < 3; a++)
        for (int q = 0; q < 4; q++)
          A->vt[PGK_T_TORS][a + 3 * ll] = A->vt[PGK_T_TORS][a + 3 * ll] + F[q][a] * P[q][ll];
  }
}

/* caltorsp / calijkl: caltors.F90:130-460 */
static int caltors(const pgko_sys *s, acc *A, int vir, const hfctx *h) {
  double pot = 0.0;
  for (int n = 0; n < s->ntors; n++) {
    int t = s->tors_type[n] - 1;
    int i = s->tors_ijkl[4 * n], j = s->tors_ijkl[4 * n + 1];
    int k = s->tors_ijkl[4 * n + 2], l = s->tors_ijkl[4 * n + 3];
    torsgeo g;
    tors_geo(s, i, j, k, l, &g);
    double per = fabs(s->tors_period[t]);
    double bh = s->tors_barhig[t], ph = s->tors_phsang[t];
    double div = (double)s->tors_divfac[t];
    double dp = bh / div * (1.0 + cos(per * g.phi - ph));
    double c = g.cphi, fa;
    int np = (int)lround(per);
    switch (np) {
      case 1: fa = 1.0; break;
      case 2: fa = 2.0 * c; break;
      case 3: fa = -4.0 * (g.sphi * g.sphi) + 3.0; break;
      case 4: fa = 4.0 * c * (2.0 * (c * c) - 1.0); break;
      case 5: fa = 4.0 * (c * c); fa = fa * (fa - 3.0) + 1.0; break;
      case 6: fa = 4.0 * (c * c); fa = 2.0 * c * (fa - 3.0) * (fa - 1.0); break;
      default: return -1;
    }
    fa = -fa * cos(ph) * bh * per / div;
    double fi[3], fj[3], fk[3], fl[3];
    tors_forces(&g, fa, fi, fj, fk, fl);
    tors_finish(s, A, vir, h, i, j, k, l, fi, fj, fk, fl, dp);
    pot = pot + dp;
  }
  A->ene[PGK_E_TORS] = pot;
  return 0;
}

/* caltorsrbp / calijklrb: caltorsrb.F90:130-459 */
static void caltorsrb(const pgko_sys *s, acc *A, int vir, const hfctx *h) {
  double pot = 0.0;
  for (int n = 0; n < s->ntorsrb; n++) {
    int t = s->rb_type[n] - 1;
    int i = s->rb_ijkl[4 * n], j = s->rb_ijkl[4 * n + 1];
    int k = s->rb_ijkl[4 * n + 2], l = s->rb_ijkl[4 * n + 3];
    torsgeo g;
    tors_geo(s, i, j, k, l, &g);
    const double *bh = s->rb_c + 6 * t;
    double b[6] = {bh[0], -bh[1], bh[2], -bh[3], bh[4], -bh[5]};
    double c = g.cphi;
    double dp = b[5], fa = 0.0;
    for (int q = 4; q >= 0; q--) {
      fa = fa * c + dp;
      dp = dp * c + b[q];
    }
    fa = -fa;
    double fi[3], fj[3], fk[3], fl[3];
    tors_forces(&g, fa, fi, fj, fk, fl);
    tors_finish(s, A, vir, h, i, j, k, l, fi, fj, fk, fl, dp);
    pot = pot + dp;
  }
  A->ene[PGK_E_TORSRB] = pot;
}

/* calnon14p: calnon14.F90:249-521; calnon14p_hf: calnon14_hf.F90:8-386 */
static void non14_pair(const pgko_sys *s, acc *A, int vir, const hfctx *h,
                       int i, int j, double dvdw, double delc, double *pe,
                       double *pv) {
  int it = s->atype[i - 1], jt = s->atype[j - 1];
  size_t tij = (size_t)(it - 1) + (size_t)s->ntype * (jt - 1);
  double r[3];
  vsub(r, POS(i), POS(j));
  minimg(s, r);
  double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  double ri2 = 1.0 / r2;
  double ri1 = sqrt(ri2);
  double pel = s->chrg[i - 1] * s->chrg[j - 1] * ri1;
  pel = pel * delc;
  double fe[3] = {pel * ri2 * r[0], pel * ri2 * r[1], pel * ri2 * r[2]};
  double rad = s->rmin[tij];
  double t2 = ri2 * (rad * rad), t6 = t2 * t2 * t2, t12 = t6 * t6;
  double e = s->eps[tij] * dvdw;
  double pvd = e * (t12 - 2.0 * t6);
  double ft = 12.0 * e * (t12 - t6) * ri2;
  double fv[3] = {ft * r[0], ft * r[1], ft * r[2]};
  add3(A->force, i, fe, 1.0);
  add3(A->force, j, fe, -1.0);
  if (vir) {
    A->ene[PGK_V_14] = A->ene[PGK_V_14] + fe[0] * r[0] + fe[1] * r[1] + fe[2] * r[2];
    tens_add(A->vt[PGK_T_14], fe, r);
  }
  *pe = *pe + pel;
  add3(A->force, i, fv, 1.0);
  add3(A->force, j, fv, -1.0);
  if (vir) {
    A->ene[PGK_V_14] = A->ene[PGK_V_14] + fv[0] * r[0] + fv[1] * r[1] + fv[2] * r[2];
    tens_add(A->vt[PGK_T_14], fv, r);
  }
  *pv = *pv + pvd;
  if (h) {
    double fej[3] = {-fe[0], -fe[1], -fe[2]}, fvj[3] = {-fv[0], -fv[1], -fv[2]};
    h->pe[PGK_HF_14][i - 1] += 0.5 * pel;
    h->pe[PGK_HF_14][j - 1] += 0.5 * pel;
    calvirihf(s, h, i, j, fe, fej, 0.5, PGK_HF_14);
    h->pe[PGK_HF_14][i - 1] += 0.5 * pvd;
    h->pe[PGK_HF_14][j - 1] += 0.5 * pvd;
    calvirihf(s, h, i, j, fv, fvj, 0.5, PGK_HF_14);
  }
}

static void calnon14(const pgko_sys *s, acc *A, int vir, const hfctx *h) {
  double pe = 0.0, pv = 0.0;
  for (int n = 0; n < s->ntors; n++) {
    int t = s->tors_type[n] - 1;
    if (!s->tors_if14[t]) continue;
    int i = s->tors_ijkl[4 * n], j = s->tors_ijkl[4 * n + 3];
    if (i == j) continue;
    double dv = 1.0 / s->tors_dvdw[t] * s->tors_ring[n];
    double de = 1.0 / s->tors_delc[t] * s->tors_ring[n];
    non14_pair(s, A, vir, h, i, j, dv, de, &pe, &pv);
  }
  for (int n = 0; n < s->ntorsrb; n++) {
    int t = s->rb_type[n] - 1;
    if (!s->rb_if14[t]) continue;
    int i = s->rb_ijkl[4 * n], j = s->rb_ijkl[4 * n + 3];
    if (i == j) continue;
    double dv = 1.0 / s->rb_dvdw[t] * s->rb_ring[n];
    double de = 1.0 / s->rb_delc[t] * s->rb_ring[n];
    non14_pair(s, A, vir, h, i, j, dv, de, &pe, &pv);
  }
  A->ene[PGK_E_ELC14] = pe;
  A->ene[PGK_E_VDW14] = pv;
}

/* ---- drivers: calforce.F90:307-1068, calforce_hf.F90:363-1451 ------------ */
static int drive(pgko_sys *s, int32_t mts_flag, const pgk_mts *m, int vir,
                 int hfmode, const hfctx *h, pgk_force_out *out) {
  int n = s->natom;
  acc A;
  memset(&A, 0, sizeof(A));
  A.force = out->force;
  memset(A.force, 0, sizeof(double) * 3 * n);
  A.fc = out->for_viri_coul;
  A.fl = out->for_viri_lj;
  if (A.fc) memset(A.fc, 0, sizeof(double) * 3 * n);
  if (A.fl) memset(A.fl, 0, sizeof(double) * 3 * n);
  if (out->ene) memcpy(A.ene, out->ene, sizeof(A.ene)); /* inout energies */
  A.ene[PGK_V_COUL] = A.ene[PGK_V_LJ] = 0.0;
  for (int k = PGK_V_BOND; k <= PGK_V_14; k++) A.ene[k] = 0.0;
  int rc = 0;
  if (!hfmode) {
    /* calforce order: ewaldr, ewk, bond, angl, tors, torsrb, 14 */
    if (m->vdw == mts_flag) { if (!s->nb_index) return PGK_ESTATE; ewaldrp(s, &A, vir, NULL); }
    if (m->ewk == mts_flag) fft_pmep(s, &A, vir);
    if (m->bond == mts_flag) calbond(s, &A, vir, NULL);
    if (m->angl == mts_flag) calangl(s, &A, vir, NULL);
    if (m->tors == mts_flag) rc = caltors(s, &A, vir, NULL);
    if (m->torsrb == mts_flag) caltorsrb(s, &A, vir, NULL);
    if (m->nonb14 == mts_flag) calnon14(s, &A, vir, NULL);
  } else {
    /* calforce_hf order: bond, angl, tors, torsrb, ewaldr, ewk, 14 */
    if (m->bond == mts_flag) calbond(s, &A, 1, h);
    if (m->angl == mts_flag) calangl(s, &A, 1, h);
    if (m->tors == mts_flag) rc = caltors(s, &A, 1, h);
    if (m->torsrb == mts_flag) caltorsrb(s, &A, 1, h);
    if (m->vdw == mts_flag) { if (!s->nb_index) return PGK_ESTATE; ewaldrp(s, &A, 1, h); }
    if (m->ewk == mts_flag) fft_pmep(s, &A, 1);
    if (m->nonb14 == mts_flag) calnon14(s, &A, 1, h);
  }
  if (rc) return PGK_EINVAL;
  A.ene[PGK_E_EWC] = s->pot_ewc;
  if (mts_flag == PGK_MTS_LONG) {
    double pi = acos(-1.0);
    double a1 = -pi / (s->alpha * s->alpha * s->box[0] * s->box[1] * s->box[2]);
    A.ene[PGK_E_EWNQ] = 0.5 * a1 * s->netchrgsq;
    if (!s->ifnetq) {
      A.ene[PGK_V_COUL] = A.ene[PGK_V_COUL] + a1 * s->netchrgsq;
      for (int d = 0; d < 3; d++)
        A.vt[PGK_T_COUL][d * 4] = A.vt[PGK_T_COUL][d * 4] + 0.5 * a1 * s->netchrgsq;
    } else {
      A.ene[PGK_V_COUL] = A.ene[PGK_V_COUL] - 0.5 * a1 * s->netchrgsq;
    }
  }
  double *e = A.ene;
  e[PGK_E_NONBON] = e[PGK_E_VDW] + e[PGK_E_ELC] + e[PGK_E_EWK] + e[PGK_E_EWC] +
                    e[PGK_E_EWNQ] + e[PGK_E_VDW14] + e[PGK_E_ELC14];
  e[PGK_E_TOT] = e[PGK_E_BOND] + e[PGK_E_ANGL] + e[PGK_E_TORS] + e[PGK_E_TORSRB] +
                 e[PGK_E_NONBON];
  if (out->ene) memcpy(out->ene, A.ene, sizeof(A.ene));
  if (out->virt) memcpy(out->virt, A.vt, sizeof(A.vt));
  return PGK_OK;
}

int pgko_calforce(pgko_sys *s, int32_t mts_flag, const pgk_mts *mts,
                  int32_t ifcalpreatom, pgk_force_out *out) {
  return drive(s, mts_flag, mts, ifcalpreatom != 0, 0, NULL, out);
}

int pgko_calforce_hf(pgko_sys *s, int32_t mts_flag, const pgk_mts *mts,
                     const pgk_hf_desc *hf, pgk_force_out *out,
                     pgk_hf_out *hfo) {
  hfctx h;
  int n = s->natom;
  h.hf = hf;
  h.natom = n;
  size_t wsz = (size_t)9 * n * hf->nregion;
  memset(hfo->pot_atm, 0, sizeof(double) * PGK_NHF * n);
  memset(hfo->viri_atm, 0, sizeof(double) * PGK_NHF * wsz);
  for (int c = 0; c < PGK_NHF; c++) {
    h.pe[c] = hfo->pot_atm + (size_t)c * n;
    h.w[c] = hfo->viri_atm + (size_t)c * wsz;
  }
  return drive(s, mts_flag, mts, 1, 1, &h, out);
}

/* ---- calheatf (calheatf.F90:8-1470) -------------------------------------
 * Volume-based (ifhfvol) and bulk forms.  Centre of atom i: the atom
 * (ATOM) or its molecule's centre of mass (MOLE), z wrapped once into
 * [0, Lz) (calheatf.F90:650-655); velocity of the centre: the atom's, or
 * the molecule's mass-weighted mean over its atoms in index order
 * (calheatf.F90:548-570).  Transport part (:615-870), atoms outer, regions
 * inner: for z1 <= z <= z2, hfkin += (m u^2 / 2) v_c and hfpot_c += e_c v_c,
 * u = v - vstrm (vstrm NULL: zero): the reference subtracts atmvel_strm from
 * atmvel (:368-369) and uses that u for the kinetic energy (:677-681) and
 * in the collision part (:1027-1330), while the transported centre
 * velocities v_c come from the lab-frame copy (atmvel_tmp, :564, 588, 644).
 * Collision part (:1027-1330), regions outer, atoms inner:
 * hfviri_c(n) += W(1,n) v1 + W(2,n) v2 + W(3,n) v3.  Only the z component
 * outside the bulk form.  Each term times 1/(Lx Ly (z2 - z1)) (:1334-1380),
 * then hftot = the terms in the reference's order (:1420-1436; the classes
 * the force path does not have are zero). */
int pgko_calheatf(const pgko_sys *s, const pgk_hf_desc *hf, const double *vel,
                  const double *vstrm, const double *mass, const double *pot_atm,
                  const double *viri_atm, double *hfout) {
  const int n = s->natom, nreg = hf->nregion;
  const int bulk = hf->mode == PGK_HF_BULK;
  if (hf->mode == PGK_HF_SURFACE) return PGK_EINVAL;
  int nmol = 0;
  for (int i = 0; i < n; i++)
    if (s->molid[i] > nmol) nmol = s->molid[i];
  double *mv = (double *)calloc((size_t)3 * (nmol + 1), sizeof(double));
  double *mm = (double *)calloc((size_t)nmol + 1, sizeof(double));
  for (int i = 0; i < n; i++) {
    const int m = s->molid[i];
    mm[m] = mm[m] + mass[i];
    for (int d = 0; d < 3; d++) mv[3 * m + d] = mv[3 * m + d] + mass[i] * vel[3 * i + d];
  }
  for (int m = 1; m <= nmol; m++) {
    const double inv = 1.0 / mm[m];
    for (int d = 0; d < 3; d++) mv[3 * m + d] = mv[3 * m + d] * inv;
  }
  const int c0 = bulk ? 0 : 2;
  double *acc = (double *)calloc((size_t)nreg * PGK_NHFJ * 3, sizeof(double));
#define ACC(r, t, c) acc[((size_t)(r) * PGK_NHFJ + (t)) * 3 + (c)]
  for (int i = 0; i < n; i++) {
    double cz;
    const double *pv;
    if (hf->hftyp[i] == PGK_HFTYP_MOLE) {
      cz = hf->molecom[3 * (s->molid[i] - 1) + 2];
      pv = mv + 3 * s->molid[i];
    } else {
      cz = s->pos[3 * i + 2];
      pv = vel + 3 * i;
    }
    if (cz < 0.0) cz = cz + s->box[2];
    else if (cz >= s->box[2]) cz = cz - s->box[2];
    double v[3];
    for (int d = 0; d < 3; d++) v[d] = vel[3 * i + d] - (vstrm ? vstrm[3 * i + d] : 0.0);
    for (int r = 0; r < nreg; r++) {
      if (!(hf->z1[r] <= cz && cz <= hf->z2[r])) continue;
      const double ke = 0.5 * mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      for (int c = c0; c < 3; c++) {
        ACC(r, PGK_HFJ_KIN, c) = ACC(r, PGK_HFJ_KIN, c) + ke * pv[c];
        for (int k = 0; k < PGK_NHF; k++)
          ACC(r, PGK_HFJ_POT + k, c) =
              ACC(r, PGK_HFJ_POT + k, c) + pot_atm[(size_t)k * n + i] * pv[c];
      }
    }
  }
  for (int r = 0; r < nreg; r++)
    for (int i = 0; i < n; i++) {
      double v[3];
      for (int d = 0; d < 3; d++) v[d] = vel[3 * i + d] - (vstrm ? vstrm[3 * i + d] : 0.0);
      for (int k = 0; k < PGK_NHF; k++) {
        /* Fortran (3,3,natom,nreg) per class: W(a,n) at a + 3n */
        const double *w = viri_atm + (((size_t)k * nreg + r) * n + i) * 9;
        for (int c = c0; c < 3; c++)
          ACC(r, PGK_HFJ_VIRI + k, c) = ACC(r, PGK_HFJ_VIRI + k, c) + w[3 * c] * v[0] +
                                        w[1 + 3 * c] * v[1] + w[2 + 3 * c] * v[2];
      }
    }
  const double area_inv = (1.0 / s->box[0]) * (1.0 / s->box[1]);
  for (int r = 0; r < nreg; r++) {
    const double vol_inv = area_inv / (hf->z2[r] - hf->z1[r]);
    for (int c = 0; c < 3; c++) {
      double tot = 0.0;
      for (int t = 0; t < PGK_HFJ_TOT; t++) {
        ACC(r, t, c) = ACC(r, t, c) * vol_inv;
        tot = t == 0 ? ACC(r, t, c) : tot + ACC(r, t, c);
      }
      ACC(r, PGK_HFJ_TOT, c) = tot;
    }
  }
#undef ACC
  memcpy(hfout, acc, sizeof(double) * (size_t)nreg * PGK_NHFJ * 3);
  free(acc);
  free(mv);
  free(mm);
  return PGK_OK;
}

/* ---- standard Ewald with per-atom terms: prepewk + ewaldkp_hf -------------
 * Wave vectors as prepewk.F90:80-150: n = (kx, ky, kz), kx in [-kmax, kmax],
 * ky in [-kmax*ry, kmax*ry], kz likewise with rz = anint(y/z ratios), kept
 * when kx^2 + ky^2/ry^2 + kz^2/rz^2 <= kmax^2 (entry 0 is n = 0 with A = 0);
 * K = n / L, A(K) = exp(-(pi/alpha)^2 |K|^2) / |K|^2.  The sum
 * (ewaldk_hf.F90:60-300): per K the structure factors S_s, S_c over atoms in
 * order; E_K = A (S_s^2 + S_c^2) / (2 pi V); per atom
 * e_i += A q_i (sin_i S_s + cos_i S_c) / (2 pi V), f_i += A (S_c sin_i -
 * S_s cos_i) K (times 2 q_i / V at the end), and for K != 0 the virial
 * tensors with k_fact(a,n) = delta_an - (2/|K|^2 + 2 pi^2/alpha^2) K_a K_n,
 * net-charge corrected when ifnetq.  Self term -alpha/sqrt(pi) q_i^2 per atom. */
int32_t pgko_ewald_waves(const double *box, int32_t kmax, double yratio, double zratio,
                         int32_t cap, double *kw, double *aw, double alpha) {
  const int ry = (int)round(yratio), rz = (int)round(zratio);
  const int ykmax = kmax * ry, zkmax = kmax * rz;
  const double pi = acos(-1.0);
  const int kmax2 = kmax * kmax;
  int nw = 1;
  if (cap < 1) return -1;
  kw[0] = kw[1] = kw[2] = 0.0;
  aw[0] = 0.0;
  for (int kx = -kmax; kx <= kmax; kx++)
    for (int ky = -ykmax; ky <= ykmax; ky++)
      for (int kz = -zkmax; kz <= zkmax; kz++) {
        if (kx == 0 && ky == 0 && kz == 0) continue;
        const double k2 = (double)(kx * kx) + (double)(ky * ky) / (double)(ry * ry) +
                          (double)(kz * kz) / (double)(rz * rz);
        if (!(k2 <= (double)kmax2)) continue;
        if (nw >= cap) return -1;
        const double n[3] = {(double)kx, (double)ky, (double)kz};
        const double kk = (n[0] / box[0]) * (n[0] / box[0]) + (n[1] / box[1]) * (n[1] / box[1]) +
                          (n[2] / box[2]) * (n[2] / box[2]);
        const double coeff = (pi / alpha) * (pi / alpha);
        aw[nw] = exp(-coeff * kk) / kk;
        for (int d = 0; d < 3; d++) kw[3 * nw + d] = n[d] / box[d];
        nw++;
      }
  return nw;
}

int pgko_ewaldk_hf(const pgko_sys *s, int32_t kmax, double yratio, double zratio,
                   int32_t ifnetq, double *force, double *pot_ewk, double *pot_viri_coul,
                   double *virt, double *pot_atm, double *viri_atm) {
  const int n = s->natom;
  const int cap = 8 * (2 * kmax + 1) * (2 * kmax + 1) * (2 * kmax + 1) *
                  (int)(round(yratio) * round(zratio)) + 8;
  double *kw = (double *)malloc(sizeof(double) * 3 * cap);
  double *aw = (double *)malloc(sizeof(double) * cap);
  const int nw = pgko_ewald_waves(s->box, kmax, yratio, zratio, cap, kw, aw, s->alpha);
  if (nw < 0) {
    free(kw);
    free(aw);
    return PGK_EINVAL;
  }
  const double pi = acos(-1.0);
  const double vol = s->box[0] * s->box[1] * s->box[2];
  const double pot_fact = 0.5 / pi / vol, pi2 = 2.0 * pi, inv_vol = 1.0 / vol;
  const double fac_vir = 2.0 * pi * pi * (1.0 / (s->alpha * s->alpha));
  const double alpha_sqrtpi = s->alpha / sqrt(pi);
  double *sn = (double *)malloc(sizeof(double) * n), *cs = (double *)malloc(sizeof(double) * n);
  double *fi = (double *)calloc((size_t)3 * n, sizeof(double));
  double ewk = 0.0, ewk_cc = 0.0;
  for (int k = 0; k < nw; k++) {
    const double *K = kw + 3 * k;
    double bsk = 0.0, bck = 0.0;
    for (int i = 0; i < n; i++) {
      const double *r = s->pos + 3 * i;
      double th = K[0] * r[0] + K[1] * r[1] + K[2] * r[2];
      th = pi2 * th;
      sn[i] = sin(th);
      cs[i] = cos(th);
      bsk = bsk + s->chrg[i] * sn[i];
      bck = bck + s->chrg[i] * cs[i];
    }
    const double pfa = pot_fact * aw[k];
    const double poti = pfa * (bsk * bsk + bck * bck);
    ewk = ewk + poti;
    for (int i = 0; i < n; i++)
      pot_atm[i] = pot_atm[i] + pfa * s->chrg[i] * (sn[i] * bsk + cs[i] * bck);
    double csk[3], cck[3];
    for (int d = 0; d < 3; d++) {
      csk[d] = aw[k] * bsk * K[d];
      cck[d] = aw[k] * bck * K[d];
    }
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) fi[3 * i + d] = fi[3 * i + d] + cck[d] * sn[i] - csk[d] * cs[i];
    if (k == 0) continue;
    const double k2 = K[0] * K[0] + K[1] * K[1] + K[2] * K[2];
    double kf[9];   /* (a, n) at a + 3n, before the per-call factor */
    for (int nn = 0; nn < 3; nn++)
      for (int a = 0; a < 3; a++) {
        kf[a + 3 * nn] = -1.0 * (2.0 / k2 + fac_vir) * K[a] * K[nn];
        if (a == nn) kf[a + 3 * nn] = kf[a + 3 * nn] + 1.0;
      }
    double fv3 = 0.0;
    if (ifnetq) {
      fv3 = pfa * s->netchrgsq;
      ewk_cc = ewk_cc + fv3 * (-1.0 + fac_vir * k2);
    }
    for (int q = 0; q < 9; q++) virt[q] = virt[q] + kf[q] * (poti - fv3);
    for (int i = 0; i < n; i++) {
      const double fv2 = pfa * s->chrg[i] * (sn[i] * bsk + cs[i] * bck);
      for (int q = 0; q < 9; q++)
        viri_atm[(size_t)9 * i + q] = viri_atm[(size_t)9 * i + q] + kf[q] * (fv2 - fv3);
    }
  }
  *pot_ewk = ewk;
  *pot_viri_coul = *pot_viri_coul + ewk + ewk_cc;
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      const double f = 2.0 * s->chrg[i] * inv_vol * fi[3 * i + d];
      force[3 * i + d] = force[3 * i + d] + f;
    }
    pot_atm[i] = pot_atm[i] - alpha_sqrtpi * s->chrg[i] * s->chrg[i];
  }
  free(sn);
  free(cs);
  free(fi);
  free(kw);
  free(aw);
  return PGK_OK;
}
