/*
 * pgk_oracle.h - CPU restatement of peachgk_md's force path (test oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load or call this code, and only as the
 * checker.  The GPU library never links or calls it.
 */
#ifndef PGK_ORACLE_H
#define PGK_ORACLE_H

#include <stdint.h>
#include "pgk_force.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pgko_sys pgko_sys;

pgko_sys *pgko_create(const pgk_system_desc *desc);
void pgko_destroy(pgko_sys *s);
void pgko_set_box(pgko_sys *s, const double *box);
void pgko_set_positions(pgko_sys *s, const double *pos);
double pgko_pot_ewc(const pgko_sys *s);

/* exclusion CSR from the bonded topology (mkexcl.F90:6-223); returns length */
int32_t pgko_mkexcl(const pgk_system_desc *desc, int32_t *index_out,
                    int32_t *list_out, int32_t capacity);

/* Verlet list: mode 0 = mklist2a, 1 = mklist_cell.  Returns nlistnball
 * (including 0 sentinels) or a negative error. */
int64_t pgko_mklist(pgko_sys *s, double rcut_book, int32_t mode);
void pgko_get_list(const pgko_sys *s, int32_t *index_out, int32_t *list_out);
void pgko_set_list(pgko_sys *s, int64_t nl, const int32_t *index_in,
                   const int32_t *list_in);

int pgko_calforce(pgko_sys *s, int32_t mts_flag, const pgk_mts *mts,
                  int32_t ifcalpreatom, pgk_force_out *out);
int pgko_calforce_hf(pgko_sys *s, int32_t mts_flag, const pgk_mts *mts,
                     const pgk_hf_desc *hf, pgk_force_out *out,
                     pgk_hf_out *hfo);

/* calheatf (calheatf.F90:8-1470), volume or bulk mode, over per-atom arrays
 * in the pgk_hf_out layout: hf[nregion*PGK_NHFJ*3] (see pgk_calheatf);
 * vstrm: the streaming velocities (atmvel_strm) or NULL. */
int pgko_calheatf(const pgko_sys *s, const pgk_hf_desc *hf, const double *vel,
                  const double *vstrm, const double *mass, const double *pot_atm,
                  const double *viri_atm, double *hfout);

/* prepewk + ewaldkp_hf (prepewk.F90:9-158, ewaldk_hf.F90:8-304): standard
 * Ewald reciprocal sum with per-atom energy and virial.  Adds into force
 * [3N], pot_viri_coul, virt [9] (Fortran (a,n)), pot_atm [N] (with the self
 * term) and viri_atm [9N] (Fortran (3,3,N)); sets pot_ewk. */
int pgko_ewaldk_hf(const pgko_sys *s, int32_t kmax, double yratio, double zratio,
                   int32_t ifnetq, double *force, double *pot_ewk, double *pot_viri_coul,
                   double *virt, double *pot_atm, double *viri_atm);
/* prepewk's wave vectors K (3 per entry) and A(K); returns their number
 * (entry 0 is K = 0) or -1 when cap is too small */
int32_t pgko_ewald_waves(const double *box, int32_t kmax, double yratio, double zratio,
                         int32_t cap, double *kw, double *aw, double alpha);

/* spline tables [16*nspltbl] (same column order as the reference harness),
 * B-spline moduli [n1+n2+n3], rkcut */
void pgko_get_tables(const pgko_sys *s, double *tbl, double *bsp,
                     double *rkcut);

/* 3-D complex FFT of the oracle (isign -1 forward, +1 backward, unscaled),
 * grid (2, n1, n2, n3) interleaved complex, x fastest. Exposed for tests.   */
void pgko_fft3d(double *grid, int32_t n1, int32_t n2, int32_t n3, int32_t isign);

#ifdef __cplusplus
}
#endif
#endif
