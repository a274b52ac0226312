// pgk_internal.h - device data layout and kernel entry points of libpgkforce.
//
// Layout in HBM (one context = one system):
//   atom order (as the host):   pos  double[3N] (Fortran atmcor(3,N))
//                               q    double[N], type int[N], molid int[N]
//   cluster order (list build): 8-atom clusters along z inside x-y columns;
//                               cl_atom int[ncl*8] (-1 = padding),
//                               slot_of int[N] (atom -> cluster slot)
//   pair list:                  ci_start int[ncl+1], cj int[npair],
//                               mask uint64[npair]  (bit ii + 8*jj: pair of
//                               i-slot ii and j-slot jj is in the list)
//   per call:                   xq  (x,y,z,q) per slot in the compute type,
//                               force accumulators (FP64, or int64 fixed point
//                               in the mixed form), reduction records.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "pgk_force.h"

#define PGK_CL 8             // atoms per cluster
#define PGK_MAXJ 2048        // max candidates / half-cluster entries per i-cluster
#define PGK_MAXORD 10        // max B-spline order
#define PGK_NPROF 8          // profiled stages: see enum PGK_PROF_* in pgk_force.h

// fixed-point scale of the deterministic accumulators (mixed form)
#define PGK_FX_SCALE 4294967296.0            // 2^32
#define PGK_FX_INV (1.0 / 4294967296.0)
#define PGK_GRID_FX_SCALE 1099511627776.0    // 2^40 for the PME charge grid
#define PGK_GRID_FX_INV (1.0 / 1099511627776.0)

// reduction record layout (doubles): one record per block of each kernel
enum {
  R_ELC = 0, R_VDW, R_VCOUL, R_VLJ,
  R_TC = 4,    // 6: xx yy zz xy xz yz  (Coulomb tensor)
  R_TL = 10,   // 6: LJ tensor
  R_EWK = 16, R_VSCA, R_TK = 18,   // 6: reciprocal tensor
  R_BOND = 24, R_ANGL, R_TORS, R_TORSRB, R_ELC14, R_VDW14,
  R_VBOND = 30, R_VANGL, R_VTORS, R_V14,
  R_TB = 34, R_TA = 40, R_TT = 46, R_T14 = 52,   // 9 each (full tensors)
  R_N = 64
};

#define HIP_CHECK(x)                                                        \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                 \
      pgk_set_error(std::string("HIP error ") + hipGetErrorString(e_) +     \
                    " at " __FILE__ ":" + std::to_string(__LINE__));        \
      return PGK_EDEVICE;                                                   \
    }                                                                       \
  } while (0)

void pgk_set_error(const std::string &msg);

// eref / (amu A^2 fs^-2): a = F/m * PGK_MD_CONV, KE = m v^2 / 2 / PGK_MD_CONV
#define PGK_MD_CONV (6.9511042e-21 / 1.66053907e-17)

// volume regions of the flux form: the pair pass keeps [4][nreg] doubles of
// LDS per wave
#define PGK_HEATF_MAXREG 1024

// device view of one calforce_hf call's heat-flux set-up and accumulators
// (layout: pe[class][atom]; w[class][region][atom][nw], nw = 9 in bulk mode,
// else 3 = the z column only)
struct HfDev {
  int mode = 0, nreg = 0, natom = 0, nw = 3;
  const double *z1 = nullptr, *z2 = nullptr;   // [nreg]
  // regions that tile [ubz0, ubz0 + nreg ubw) in equal bins (ubin = 1): a
  // segment's regions are then found from its ends instead of by a walk;
  // tiled = 1 when moreover each region ends exactly where the next begins
  // (z2[r] == z1[r + 1]) and z1 increases strictly: a segment's regions
  // strictly between its end regions then take its full width
  int ubin = 0, tiled = 0;
  double ubz0 = 0.0, ubw = 1.0, ubwi = 1.0;
  const int32_t *hftyp = nullptr;              // [natom] 1 = ATOM, 2 = MOLE
  const int32_t *molid = nullptr;              // [natom] 1-based
  const double *molecom = nullptr;             // [3*nmol]
  const double *pos = nullptr;                 // [3*natom]
  double L[3] = {0, 0, 0}, Li[3] = {0, 0, 0};
  void *pe = nullptr;
  void *w = nullptr;
  // flux form (pgk_calforce_heatf): instead of the W arrays, the collision
  // part of calheatf, sum_a W(a,n,i,r) v_i(a), is accumulated directly
  int flux = 0;
  const double *vel = nullptr;                 // [3*natom] collision-part velocities
  // the transport part's velocities (set by the flux form's callers); with
  // one bulk region over the whole box, atom centres only and tvel == vel
  // (direct_ok and the pointers equal) the pair pass sums the pair
  // energies' transport share itself (direct_tr) instead of per-atom energies
  const double *tvel = nullptr;
  int direct_ok = 0, direct_tr = 0;
  double *jcol = nullptr;   // [PGK_NHF][nreg][3] sums, heat-flux accumulator type (int64 fixed point in the mixed form)
  double *jpart = nullptr;                     // pair pass: per-block records
};

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
};

struct pgk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // the reciprocal-space chain runs on stream2 beside the pair kernel when
  // nothing needs the stages one after the other (fork / join events)
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool pme_overlap = true;
  int pair_cap_blocks = 0;         // >0: pair blocks per CU while the chain runs beside
  int lds_per_cu = 0;              // bytes of LDS per CU
  int precision = PGK_PREC_FP64;

  // system (host copies)
  int natom = 0, ntype = 0;
  std::vector<double> charge, lj_eps, lj_rmin;
  std::vector<int32_t> atype, molid, inttyp, excl_index, excl_list;
  std::vector<int32_t> bond_ij, bond_type, angl_ijk, angl_type;
  std::vector<double> bond_k, bond_r0, angl_k, angl_t0;
  std::vector<int32_t> tors_ijkl, tors_type, tors_divfac, tors_if14;
  std::vector<double> tors_ring, tors_barhig, tors_phsang, tors_period, tors_dvdw,
      tors_delc;
  std::vector<int32_t> rb_ijkl, rb_type, rb_if14;
  std::vector<double> rb_ring, rb_c, rb_dvdw, rb_delc;
  double alpha = 0, rrcut = 0, rcut = 0;
  double rrsq_cut = 0;   // largest r^2 with sqrt(r^2) <= rrcut (correctly rounded)
  int nspl = 0, nfft[3] = {0, 0, 0}, order = 4, ifnetq = 0;
  double box[3] = {0, 0, 0};
  double netchrgsq = 0, pot_ewc = 0, rkcut = 0, spl_int = 0;
  std::vector<double> bsp_mod[3];
  bool have_pos = false, have_list = false;
  double list_box[3] = {0, 0, 0};

  // device: atom order
  double *d_pos = nullptr;      // 3N
  double *d_q = nullptr;        // N
  int32_t *d_type = nullptr;    // N
  int32_t *d_molid = nullptr;   // N
  double *d_eps = nullptr, *d_rmin = nullptr;   // ntype^2 (column-major)
  int32_t *d_inttyp = nullptr;
  int32_t *d_excl_index = nullptr, *d_excl_list = nullptr;
  int32_t nexclpair = 0;
  int32_t *d_exclpair = nullptr;   // 2 * nexclpair, 0-based (i<j)
  int32_t *d_xs_start = nullptr;   // N+1: symmetric exclusion CSR (both directions)
  int32_t *d_xs_j = nullptr;       // 2 * nexclpair partners, 0-based
  int32_t *d_xl_start = nullptr;   // N+1: the same over entries j > i only (list search)
  int32_t *d_xl_j = nullptr;
  double *d_spl = nullptr;         // nspl * 8 (real a,b,c,d | realpot a,b,c,d)
  double *d_splx = nullptr;        // nspl * 8 (excs | excspot)
  float *d_splf = nullptr;         // nspl * 8 float copy (real | realpot)
  double *d_ljp64 = nullptr;       // ntype^2 x (eps, rmin^2)
  float *d_ljp32 = nullptr;
  double *d_bsp = nullptr;         // n1+n2+n3

  // bonded (device, 0-based indices)
  int32_t *d_bond = nullptr;       // 3 per bond: i j type
  int32_t *d_angl = nullptr;       // 4 per angle
  int32_t *d_tors = nullptr;       // 5 per torsion
  int32_t *d_rb = nullptr;         // 5 per RB torsion
  int32_t *d_p14 = nullptr;        // 2 per 1-4 pair
  double *d_p14f = nullptr;        // 2 per 1-4 pair: vdw factor, elc factor
  double *d_bpar = nullptr;        // bond k, r0 | angle k, t0 | tors ... packed
  int n14 = 0;
  // bonds / angles whose force constant passes the reference's skip test
  // (calbond.F90:70, calangl.F90): terms below it never contribute
  int nbond_act = 0, nangl_act = 0;
  int off_bk = 0, off_br = 0, off_ak = 0, off_at = 0, off_tb = 0, off_tp = 0,
      off_tper = 0, off_tdiv = 0, off_rbc = 0;

  // clusters / list
  int ncol[2] = {1, 1}, nzb = 1, nbin = 0;
  int ncl = 0;                     // number of clusters
  int32_t *d_bin_count = nullptr, *d_bin_start = nullptr, *d_atom_bin = nullptr,
          *d_atom_slot = nullptr, *d_sorted = nullptr;
  int32_t *d_col_ncl = nullptr, *d_col_clstart = nullptr;  // per column
  int32_t *d_cl_atom = nullptr;    // ncl*8
  int32_t *d_cl_col = nullptr;     // ncl: column of each cluster
  int32_t *d_slot_of = nullptr;    // N
  float *d_bbox = nullptr;         // ncl * 8: lo xyz, hi xyz (wrapped frame)
  double4 *d_posc = nullptr;       // ncl*8: FP64 position + atom index per slot
  int32_t *d_tmp_cj = nullptr;     // ncl * PGK_MAXJ
  uint32_t *d_tmp_mask = nullptr;
  int32_t *d_ci_count = nullptr, *d_ci_start = nullptr;
  // the pair kernel's work items in its chunked form (64 list entries of one
  // cluster each, coded cluster * 32 + chunk) and their count
  int32_t *d_wcount = nullptr, *d_wstart = nullptr, *d_witem = nullptr;
  int nwi = 0;
  int64_t pair_chunk_below = PGK_TUNE_DEFAULT_CHUNK_BELOW;   // chunked form below this ncl
  int pair_cap_default = 2;        // pair blocks per CU beside the chain (0: no cap)
  int pair_cap_chunked = 3;        // the same for the chunked form
  int32_t *d_cj = nullptr;         // entries: half-cluster index 2 cj + h
  uint32_t *d_mask = nullptr;      // entries: bit ii + 8 jq for i-slot ii, j-slot 4 h + jq
  int64_t npair_cl = 0, cap_pair_cl = 0;
  int32_t *d_list_meta = nullptr;  // cluster count, overflow flags, entry count
  int32_t *h_list_meta = nullptr;  // pinned copy read once per rebuild
  int64_t npairs = 0;              // atom pairs in the list
  int32_t *d_overflow = nullptr;
  void *d_scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
  int ncl_cap = 0, nbin_cap = 0, ncolumn_cap = 0;

  // per call
  void *d_xq = nullptr;            // ncl*8 of double4 or float4
  int32_t *d_ct = nullptr;         // ncl*8 types (0 for padding)
  void *d_fcl = nullptr;           // ncl*8*3 force accumulators (slot order)
  void *d_fat = nullptr;           // N*3 force accumulators (atom order)
  void *d_fcl_c = nullptr;         // slot-order Coulomb part (for_viri_coul)
  void *d_fat_c = nullptr;         // atom-order Coulomb part
  double *d_force = nullptr;       // N*3 final force (double)
  double *d_fvc = nullptr, *d_fvl = nullptr;   // per-class forces (optional)
  double *d_red = nullptr;         // reduction records
  int red_cap = 0;
  double *d_sum = nullptr;         // R_N final sums
  double *h_sum = nullptr;         // pinned R_N

  // PME
  void *d_grid = nullptr;          // charge grid (n1*n2*n3) in the compute type, written
                                   // point by point by the tile spread
  void *d_pgrid = nullptr;         // potential grid (n1*n2*n3) in the compute type
  void *d_cgrid = nullptr;         // complex half grid ((n1/2+1)*n2*n3)
  void *d_theta = nullptr;         // B-spline weights + base indices per atom
  size_t theta_bytes = 0;
  void *d_tile = nullptr;          // spreading bins: counts, starts, ranks, items, scan space
  size_t tile_bytes = 0;
  int tile_ntile = 0;
  double *d_tw = nullptr;          // twiddles: 3 dims, each n complex (FP64)
  float *d_twf = nullptr;
  int tw_off[3] = {0, 0, 0};

  // heat flux (calforce_hf): inputs, accumulators and export staging
  HfDev hfdev;
  double *d_hf_z = nullptr;        // z1 | z2
  int32_t *d_hf_typ = nullptr;     // N
  double *d_hf_com = nullptr;      // 3*nmol
  void *d_hf_pe = nullptr, *d_hf_w = nullptr;
  double *d_hf_pe_out = nullptr, *d_hf_w_out = nullptr;
  size_t hf_z_cap = 0, hf_com_cap = 0, hf_w_cap = 0, hf_wout_cap = 0;
  bool hf_ready = false;           // a pgk_calforce_hf call left its arrays on the device
  bool hf_export = true;           // run_calforce writes the caller-layout HF arrays
  bool hf_pairs_only = false;      // run_calforce: pair heat-flux pass without pair forces
  // W-array pair pass (k_hf_w.hip): per-slot FP64 positions, centres and
  // molecule ids, and the transposed list (entries grouped by half-cluster)
  double4 *d_hfw_xq = nullptr, *d_hfw_cen = nullptr;
  int32_t *d_hfw_mol = nullptr;
  size_t hfw_xq_bytes = 0, hfw_cen_bytes = 0, hfw_mol_bytes = 0;
  double4 *d_hfw_v = nullptr;   // per slot: collision-part velocity (fused flux pass)
  size_t hfw_v_bytes = 0;
  int32_t *d_tl_start = nullptr, *d_tl_src = nullptr;
  uint32_t *d_tl_mask = nullptr;
  size_t tl_start_bytes = 0, tl_src_bytes = 0, tl_mask_bytes = 0;
  int64_t list_serial = 0, tl_serial = -1;   // list builds; the build the transpose is of
  bool pme_grid_ready = false;     // d_pgrid holds the potential of the current positions
  // per-atom reciprocal virial (pgk_recip_virial_atom): one component's
  // potential grid at a time, and the six per-atom components
  void *d_vgrid = nullptr;
  size_t vgrid_bytes = 0;
  double *d_vir6 = nullptr;
  size_t vir6_bytes = 0;
  // calheatf (pgk_calheatf): velocities, masses, molecule lists, partial records
  double *d_hj_vel = nullptr, *d_hj_mass = nullptr, *d_hj_molevel = nullptr;
  double *d_hj_cz = nullptr;   // centre z per atom (transport part, many regions)
  size_t hj_cz_cap = 0;
  double *d_hj_part = nullptr, *d_hj_out = nullptr;
  int32_t *d_hj_mol_start = nullptr, *d_hj_mol_atoms = nullptr;   // atoms of each molecule
  double *d_hj_jcol = nullptr, *d_hj_jpart = nullptr;   // flux form sums
  // heat-flux time series of the MD loop (pgk_md_heatf)
  int md_hf_interval = 0, md_hf_rows = 0;
  std::vector<double> md_hf_z1, md_hf_z2;
  std::vector<int32_t> md_hf_typ;
  int md_hf_mode = 0;
  bool md_hf_mole = false;         // some atoms are MOLE centres (device calcom)
  double *d_md_hfseries = nullptr, *d_md_hfmass = nullptr, *d_hj_fscratch = nullptr;
  size_t md_hfseries_cap = 0;
  std::vector<double> md_mass;     // masses (amu) given to pgk_md_setup
  // RESPA (pgk_md_respa): inner steps per outer step and the MTS level of
  // each term; the short force lives in d_md_fshort
  int md_nshort = 1;
  pgk_mts md_mts = {PGK_MTS_LONG, PGK_MTS_LONG, PGK_MTS_LONG, PGK_MTS_LONG,
                    PGK_MTS_LONG, PGK_MTS_LONG, PGK_MTS_LONG};
  double *d_md_fshort = nullptr;
  double md_e_short[PGK_NENE] = {};   // the short call's energy record at a sample
  size_t hj_jcol_cap = 0, hj_jpart_cap = 0;
  int hj_nmol = 0;
  bool hj_have_mass = false;
  size_t hj_part_cap = 0, hj_out_cap = 0;

  // device-resident MD step (velocity Verlet + RATTLE)
  bool md_ready = false;
  int md_ngroup = 0, md_ncons = 0;
  double md_dt = 1.0, md_tol = 1e-10;
  int32_t *d_md_gstart = nullptr, *d_md_cij = nullptr, *d_md_fail = nullptr;
  int32_t *d_md_small = nullptr;
  bool md_all_tri = false;         // every group a triangle: fused half-step kernels
  bool md_graphs = false;          // pgk_md_run: replay plain steps from a captured graph
  bool md_batch_graphs = false;    // pgk_md_run_batch: the same, per replica
  int md_nfree = 0;                // atoms in no constraint group
  int32_t *d_md_free = nullptr;
  double *d_md_d2 = nullptr, *d_md_minv = nullptr, *d_md_vel = nullptr, *d_md_pos_old = nullptr;
  double *d_md_ke = nullptr;       // kinetic-energy block records
  double *d_md_kenhc = nullptr;    // the chain's records (atoms outside Langevin regions)
  // Nose-Hoover chain (pgk_md_nhc); off when md_nhc_m == 0
  int md_nhc_m = 0, md_nhc_nyosh = 1, md_nhc_next = 1;
  double md_nhc_gkt = 0.0, md_nhc_gnkt = 0.0;
  double md_nhc_ys[5] = {};
  double *d_md_nhc = nullptr;      // [3m+2]: xlogs, vlogs, qmass, exp(-dt vxi1/2), bbv
  // region Langevin thermostat (pgk_md_langevin); off when md_lt_nreg == 0.
  // Registration units: a molecule by its centre of mass (ucom 1) or one
  // atom (ucom 0), CSR ustart over uatom.
  int md_lt_nreg = 0, md_lt_nunit = 0;
  double md_lt_lo[3][PGK_LT_MAXREG] = {}, md_lt_hi[3][PGK_LT_MAXREG] = {};
  double md_lt_temp[PGK_LT_MAXREG] = {}, md_lt_damp[PGK_LT_MAXREG] = {};
  uint64_t md_lt_seed = 0, md_lt_draw = 0;   // draw: steps run since set-up
  int md_lt_nstrm = 0;             // streaming-velocity slabs
  double md_lt_sz[2][PGK_LT_MAXREG] = {}, md_lt_sv[3][PGK_LT_MAXREG] = {};
  int32_t *d_lt_sreg = nullptr;    // streaming slab of each atom (-1 none)
  double *d_md_vstrm = nullptr;    // [3*natom] streaming velocity of each atom
  // atmvel_strm of the standalone heat-flux calls (pgk_set_atmvel_strm)
  double *d_hf_vstrm = nullptr;
  bool hf_vstrm_on = false;
  double *d_hj_vpec = nullptr;     // velocities minus the streaming ones
  size_t hj_vpec_cap = 0;
  int32_t *d_lt_ustart = nullptr, *d_lt_uatom = nullptr, *d_lt_ucom = nullptr;
  int32_t *d_lt_reg = nullptr, *d_lt_cnt = nullptr;   // region of each atom (-1 none)
  double *d_lt_mass = nullptr, *d_lt_frand = nullptr;
  double *d_lt_det = nullptr;      // [block][PGK_LT_MAXREG] energy records

  // reduction segment bookkeeping for the current call
  int red_used = 0;
  // accumulators (force sums, charge grid) are cleared by the kernels that
  // consume them; set when a call did not run to completion
  bool acc_dirty = true;

  // optional per-stage timing: every stage instance records a begin/end pair
  // of HIP events from a pool; pairs are read at the next synchronisation
  bool prof = false;
  uint32_t prof_mask = 0xffffffffu;   // stages that record events
  std::vector<hipEvent_t> pev;     // 2 per pending stage instance
  std::vector<int> pstage;
  int pn = 0;                      // pending stage instances
  int popen[PGK_NPROF] = {};
  double prof_ms[PGK_NPROF] = {};
  int64_t prof_calls[PGK_NPROF] = {};

  // the last calforce call, finalised by pgk_energies (or at once when the
  // call waits): its reduction segments and the inputs of the energy record
  int last_nseg = 0;
  int last_seg_kind[16] = {};
  int seg_up[3][32] = {};          // segment tables currently on the device (3 slots)
  int seg_up_n[3] = {-1, -1, -1};
  int32_t last_mts_flag = 0;
  pgk_mts last_mts = {};
  bool last_pending = false;
};

// stage markers (record before/after a stage when profiling is on)
void pgk_prof_begin(pgk_ctx *c, int k);
void pgk_prof_end(pgk_ctx *c, int k);
void pgk_prof_collect(pgk_ctx *c);

// ---- kernels / launchers (return hipError_t of the launch) ----------------
int pgk_list_build(pgk_ctx *c, double rcut_book);
int pgk_list_count_pairs(pgk_ctx *c);   // fills c->npairs (counted on request)
int pgk_list_count_pairs(pgk_ctx *c);   // fills c->npairs (counted lazily)
int pgk_pairs_launch(pgk_ctx *c, bool want_fsplit, bool energies, int *red_off, int *red_n);
int pgk_combine_launch(pgk_ctx *c, bool do_pair, bool split, double *fout, double *fvc,
                       double *fvl);
int pgk_excl_launch(pgk_ctx *c, bool split, int *off, int *nrec);
int pgk_pme_launch(pgk_ctx *c, bool want_fsplit, int *red_off, int *red_n);
int pgk_pme_prepare(pgk_ctx *c);
int pgk_pme_prep(pgk_ctx *c, bool pack);
int pgk_bonded_launch2(pgk_ctx *c, const pgk_mts *m, int mts_flag, int *seg_off, int *seg_n,
                       int *seg_kind, int *nseg, bool hf);
int pgk_hf_launch_pairs(pgk_ctx *c, bool do_pairs);
// regions the W-array pair pass holds per wave in LDS (volume / surface forms)
#define PGK_HF_W_MAXREG 64
int pgk_hf_w_launch(pgk_ctx *c);
int pgk_hf_export(pgk_ctx *c);
// d_vstrm: streaming velocities [3*natom] (kinetic energies and collision
// part relative to them), or null; the rows go to out (nreg x PGK_NHFJ x 3
// doubles on the device), or to c->d_hj_out when out is null
int pgk_heatf_launch(pgk_ctx *c, const double *d_vel, const double *d_mass,
                     const double *d_vstrm, double *out = nullptr);
// zero n0, n1 and n2 8-byte words at p0, p1 and p2 (a null pointer with a
// zero count is skipped) in one kernel launch on st
int pgk_zero3_launch(hipStream_t st, void *p0, size_t n0, void *p1, size_t n1, void *p2,
                     size_t n2);
// per cluster slot: position and charge, heat-flux centre, molecule id and
// (vel4 non-null) the collision-part velocity, for the heat-flux pair passes
__global__ void k_hf_pack(int nslot, const int32_t *__restrict__ cl_atom, HfDev h,
                          const double *__restrict__ q, double4 *__restrict__ xq,
                          double4 *__restrict__ cen, int32_t *__restrict__ mol,
                          double4 *__restrict__ vel4);
int pgk_heatf_collision_vel(pgk_ctx *c, const double *d_vel, const double *d_vstrm,
                            const double **out);
// molecule centres of mass from d_pos (calcom.F90:1-147) into d_com [3*nmol]
int pgk_calcom_launch(pgk_ctx *c, const double *d_mass, double *d_com);
int pgk_pme_atom_energy(pgk_ctx *c, double *d_e);
// atoms the allocation-free entry points' fixed scratch holds (8 MiB of doubles)
#define PGK_NOALLOC_MAX_ATOMS (1 << 20)
int pgk_noalloc_init_device(int dev);     // once per device, that device current
int pgk_pme_atom_virial(pgk_ctx *c, double *d_w);   // (3,3,natom) into d_w
// grow a device buffer to at least `need` bytes (contents not kept)
extern "C" int grow_dev(void **p, size_t *cap, size_t need);
int pgk_md_kick_drift(pgk_ctx *c, const double *f, double dt, const double *fl, double dtl);
int pgk_md_kick(pgk_ctx *c, const double *f, double dt, const double *fl, double dtl);
int pgk_md_rattle_v(pgk_ctx *c);
int pgk_md_ke(pgk_ctx *c, double *rec, bool nhc = false);
int pgk_md_nhc_half(pgk_ctx *c);
int pgk_md_lt_register(pgk_ctx *c);      // assign atoms to the Langevin regions
int pgk_md_strm_assign(pgk_ctx *c);      // streaming velocities at the current positions
int pgk_md_lt_half(pgk_ctx *c, bool first);
int pgk_md_lt_count(pgk_ctx *c);         // natom_lt into d_lt_cnt

// raise kernel f's dynamic LDS limit to at least lds bytes: the runtime call
// is made once per kernel, device and larger size (it costs host time on
// every launch otherwise)
int pgk_lds_limit(const void *f, size_t lds);

static inline int pgk_grid_blocks(int64_t n, int bs) {
  return (int)((n + bs - 1) / bs);
}

// the pair passes over the list (the force pass and the fused heat-flux pass)
// run one wave per 64 entries of a cluster instead of one wave per cluster
// when the system has too few clusters to fill the device
static inline bool pgk_pair_chunked(const pgk_ctx *c) {
  return (int64_t)c->ncl < c->pair_chunk_below;
}

// waves of a pass over the current list in the form pgk_pair_chunked picks
static inline int64_t pgk_pair_waves(const pgk_ctx *c) {
  return pgk_pair_chunked(c) ? (int64_t)c->nwi : (int64_t)c->ncl;
}
