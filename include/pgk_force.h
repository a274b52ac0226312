/*
 * pgk_force.h - C ABI of the MI355X force-evaluation library (libpgkforce.so),
 * a drop-in for peachgk_md's force path.
 *
 * The reference evaluates forces in two Fortran drivers that read all system
 * data from module globals (md_global.inc:143-470) and take the box, cutoffs
 * and flags as arguments:
 *   calforce     peachgk-md/calforce.F90:8-1073      (interface_mdtech.F90:358)
 *   calforce_hf  peachgk-md/calforce_hf.F90:8-1451   (interface_mdtech.F90:609)
 * and rebuild their Verlet list with
 *   mklist2a     peachgk-md/mklist.F90:8-234
 *   mklist_cell  peachgk-md/mklist_cell.F90:8-352
 * This header replaces those entry points with plain-pointer C functions that a
 * Fortran host binds with iso_c_binding (see the pgk_force_mod module shipped
 * beside the library).  All arrays are host arrays in the reference's own
 * layout and units: positions and forces are Fortran (3,natom) column-major
 * doubles, atom indices in topology arrays are 1-based, lengths are in the
 * reference's non-dimensional unit (1 A), energies in eref, and charges are
 * already multiplied by 1/sqrt(4*pi*eps0) (prepmd.F90:239-242).
 *
 * Every function returns PGK_OK (0) or a negative PGK_E* code; the message of
 * the last failure on the calling thread is returned by pgk_last_error().
 */
#ifndef PGK_FORCE_H
#define PGK_FORCE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGK_ABI_VERSION 1

/* ---- status codes -------------------------------------------------------- */
enum {
  PGK_OK = 0,
  PGK_EINVAL = -1,     /* bad argument (size, index, flag)                    */
  PGK_EDEVICE = -2,    /* HIP runtime failure or no usable gfx950 device      */
  PGK_ENOMEM = -3,     /* device allocation failed                            */
  PGK_ESTATE = -4,     /* call out of order (e.g. no list built yet)          */
  PGK_EOVERFLOW = -5   /* list or cell capacity exceeded (mklist_cell.F90:121-145,342-346) */
};

/* ---- arithmetic forms ---------------------------------------------------- */
enum {
  PGK_PREC_FP64 = 0,   /* FP64 throughout, as the reference                    */
  PGK_PREC_MIXED = 1,  /* FP32 pair/spline/FFT arithmetic, 64-bit fixed-point
                          force/virial/heat-flux sums, FP64 reductions of
                          energies and the Ewald self/net-charge terms          */
  PGK_PREC_FP32 = 2    /* FP32 everywhere inside the force evaluation         */
};

/* ---- MTS levels (md_global.inc:138-140) ---------------------------------- */
enum { PGK_MTS_LONG = 1, PGK_MTS_MED = 2, PGK_MTS_SHORT = 3 };

/* ---- heat-flux virial modes (calvirihf.F90:8-405) ------------------------ */
enum {
  PGK_HF_VOLUME = 0,   /* default build: z-fragment weighted column 3 (ifhfvol) */
  PGK_HF_SURFACE = 1,  /* ifhfvol = .false.: plane-crossing (method of planes)  */
  PGK_HF_BULK = 2      /* -D_HF_BULK: full 3x3, minimum image on all axes       */
};
enum { PGK_HFTYP_ATOM = 1, PGK_HFTYP_MOLE = 2 };   /* md_global.inc:142-143 */

/* ---- energy record: index into pgk_force_out.ene ------------------------- */
enum {
  PGK_E_TOT = 0, PGK_E_NONBON, PGK_E_VDW, PGK_E_ELC, PGK_E_EWK, PGK_E_EWC,
  PGK_E_EWNQ, PGK_E_VDW14, PGK_E_ELC14, PGK_E_BOND, PGK_E_ANGL, PGK_E_TORS,
  PGK_E_TORSRB,
  PGK_V_COUL,          /* pot_viri_coul */
  PGK_V_LJ,            /* pot_viri_lj   */
  PGK_V_BOND,          /* atm_viri_bond */
  PGK_V_ANGL,          /* atm_viri_angl */
  PGK_V_TORS,          /* atm_viri_tors (periodic + RB) */
  PGK_V_14,            /* atm_viri_14   */
  PGK_NENE
};
/* ---- virial tensors: 3x3 blocks (Fortran order, (force comp, r comp)) ----- */
enum {
  PGK_T_COUL = 0,      /* pot_virit_coul (real + excluded + reciprocal + netq) */
  PGK_T_LJ,            /* pot_virit_lj   */
  PGK_T_BOND,          /* atm_virit_bond */
  PGK_T_ANGL,          /* atm_virit_angl */
  PGK_T_TORS,          /* atm_virit_tors */
  PGK_T_14,            /* atm_virit_14   */
  PGK_NTENS
};
/* ---- per-atom heat-flux classes (calforce_hf.F90:300-330) ---------------- */
enum {
  PGK_HF_BO = 0,       /* pot_bo_atm   / viribot_atm   */
  PGK_HF_AN,           /* pot_an_atm   / viriant_atm   */
  PGK_HF_TO,           /* pot_to_atm   / viritot_atm   */
  PGK_HF_14,           /* pot_14_atm   / viri14t_atm   */
  PGK_HF_ELIN,         /* pot_elin_atm / virielint_atm */
  PGK_HF_LJIN,         /* pot_ljin_atm / viriljint_atm */
  PGK_HF_ELC,          /* pot_elc_atm  / virielct_atm  */
  PGK_HF_VDW,          /* pot_vdw_atm  / viriljt_atm   */
  PGK_NHF
};

/*
 * System description: the part of md_global / spme_global the force path
 * reads (md_global.inc:143-470, spme_global.inc:24-43).  Copied at create
 * time; the caller keeps ownership of every array.
 */
typedef struct pgk_system_desc {
  int32_t natom;
  int32_t ntype;               /* LJ tables are ntype x ntype                 */
  const double *charge;        /* [natom] atmchrg (scaled)                     */
  const int32_t *atype;        /* [natom] atmindex, 1..ntype                   */
  const int32_t *molid;        /* [natom] irmolept_list, 1-based               */
  const double *lj_eps;        /* [ntype*ntype] vdw_welij (column-major)       */
  const double *lj_rmin;       /* [ntype*ntype] vdw_radij (R_min)              */
  const int32_t *inttyp;       /* [ntype*ntype] inter_inttyp (1 = INTTYPE_VDW) */
  /* exclusion CSR as left by mkexcl (mkexcl.F90:193-222)                      */
  int32_t nexcl;               /* length of excl_list                          */
  const int32_t *excl_index;   /* [natom+1] 1-based start of each atom's run   */
  const int32_t *excl_list;    /* [nexcl] partners j > i, or a single 0        */
  /* harmonic bonds (calbond.F90)                                              */
  int32_t nbond, nbondtyp;
  const int32_t *bond_ij;      /* [2*nbond] (i,j) pairs, 1-based               */
  const int32_t *bond_type;    /* [nbond] 1-based                              */
  const double *bond_k, *bond_r0;            /* [nbondtyp]                     */
  /* harmonic angles (calangl.F90)                                             */
  int32_t nangl, nangltyp;
  const int32_t *angl_ijk;     /* [3*nangl]                                    */
  const int32_t *angl_type;
  const double *angl_k, *angl_theta0;        /* [nangltyp], theta0 in rad      */
  /* periodic torsions (caltors.F90) and their 1-4 pairs (calnon14.F90)        */
  int32_t ntors, ntorstyp;
  const int32_t *tors_ijkl;    /* [4*ntors]                                    */
  const int32_t *tors_type;
  const double *tors_ring14;   /* [ntors] ring14                               */
  const int32_t *tors_divfac;  /* [ntorstyp] para_divfac                       */
  const double *tors_barhig, *tors_phsang, *tors_period;   /* [ntorstyp]       */
  const double *tors_divfac_vdw, *tors_divfac_elc;         /* [ntorstyp]       */
  const int32_t *tors_if14;    /* [ntorstyp] ifcalnonb14 (0/1)                 */
  /* Ryckaert-Bellemans torsions (caltorsrb.F90) and their 1-4 pairs           */
  int32_t ntorsrb, ntorsrbtyp;
  const int32_t *torsrb_ijkl;
  const int32_t *torsrb_type;
  const double *torsrb_ring14;
  const double *torsrb_c;      /* [6*ntorsrbtyp] para_barhigrb(1:6,t)          */
  const double *torsrb_divfac_vdw, *torsrb_divfac_elc;
  const int32_t *torsrb_if14;
  /* Ewald / SPME parameters (already non-dimensional)                        */
  double alpha;                /* Ewald alpha [1/A]                            */
  double rrcut;                /* Ewald real-space cutoff                      */
  double rcut;                 /* LJ cutoff                                    */
  int32_t nspltbl;             /* spline table points (spline_interp.F90:74)   */
  int32_t nfft1, nfft2, nfft3; /* SPME grid                                    */
  int32_t pme_order;           /* B-spline order, 3..10                        */
  int32_t ifnetqcorrp;         /* net-charge correction for pressure (0/1)     */
  const double *box;           /* [3] initial box: fixes rkcut (erf_corr_cutoff.F90:28-62) */
} pgk_system_desc;

/* MTS level of each force term, as the mts_* arguments of calforce. A term is
 * evaluated when its level equals the mts_flag of the call; 0 disables it.  */
typedef struct pgk_mts {
  int32_t bond, angl, tors, torsrb, vdw, ewk, nonb14;
} pgk_mts;

/* Outputs of one calforce call.  force and ene are required; the rest may be
 * NULL.  Semantics follow calforce.F90:307-345 (everything zeroed on entry).  */
typedef struct pgk_force_out {
  double *force;               /* [3*natom]                                    */
  double *ene;                 /* [PGK_NENE]                                   */
  double *virt;                /* [9*PGK_NTENS]                                */
  double *for_viri_coul;       /* [3*natom] Coulomb part of the force          */
  double *for_viri_lj;         /* [3*natom] LJ part of the force               */
} pgk_force_out;

/* Heat-flux region set-up: the extra arguments of calforce_hf. */
typedef struct pgk_hf_desc {
  int32_t mode;                /* PGK_HF_VOLUME / _SURFACE / _BULK              */
  int32_t nregion;
  const double *z1, *z2;       /* [nregion] hfzpos1 / hfzpos2                  */
  const int32_t *hftyp;        /* [natom] hftyp_atm                            */
  int32_t nmol;
  const double *molecom;       /* [3*nmol] molecule centres of mass (calcom)   */
} pgk_hf_desc;

/* ---- heat-flux terms of calheatf (calheatf.F90:1334-1463): index t of
 * hf[(r*PGK_NHFJ + t)*3 + comp] --------------------------------------------- */
enum {
  PGK_HFJ_KIN = 0,                 /* hfkin_atm                                 */
  PGK_HFJ_POT = 1,                 /* + PGK_HF_* class: hfpot*_atm              */
  PGK_HFJ_VIRI = 1 + PGK_NHF,      /* + PGK_HF_* class: hfviri*_atm             */
  PGK_HFJ_TOT = 1 + 2 * PGK_NHF,   /* hftot_atm                                 */
  PGK_NHFJ
};

/* Per-atom heat-flux outputs: PGK_NHF classes, each natom energies and a
 * Fortran (3,3,natom,nregion) virial array.  Either pointer (or hfo itself
 * in pgk_calforce_hf) may be NULL: that array then stays on the device only,
 * for pgk_calheatf.                                                          */
typedef struct pgk_hf_out {
  double *pot_atm;             /* [PGK_NHF*natom]                               */
  double *viri_atm;            /* [PGK_NHF*9*natom*nregion]                     */
} pgk_hf_out;

typedef struct pgk_ctx pgk_ctx;

/* Library version and the last error message of the calling thread. */
int32_t pgk_abi_version(void);
const char *pgk_last_error(void);

/* Create a context on the current HIP device: copies the description, builds
 * the erfc spline tables (spline_interp.F90:14-136), the B-spline moduli and
 * rkcut (fft_pme_init.F90:14-182, erf_corr_cutoff.F90:28-62) and the Ewald
 * self energy.  Replaces the module set-up done by main_md.F90:1101-1187.   */
int pgk_create(const pgk_system_desc *desc, int32_t precision, pgk_ctx **out);
int pgk_destroy(pgk_ctx *ctx);

/* Box lengths [3] (xcel, ycel, zcel).  rkcut keeps its creation-time value,
 * as in the reference.                                                        */
int pgk_set_box(pgk_ctx *ctx, const double *box);
/* Copy positions (Fortran atmcor(3,natom)) from host memory. */
int pgk_set_positions(pgk_ctx *ctx, const double *pos);
/* Same from device memory (no host round trip). */
int pgk_set_positions_device(pgk_ctx *ctx, const double *d_pos);

/* Verlet list rebuild: the pair set of mklist2a / mklist_cell (i<j, r^2 <
 * rcut_book^2 in FP64 minimum image, exclusions removed), stored as cluster
 * pairs with per-pair masks.  Replaces mklist2a / mklist_cell.              */
int pgk_build_list(pgk_ctx *ctx, double rcut_book);
/* Number of listed pairs (nlistnball without the 0 sentinels). */
int pgk_list_npairs(pgk_ctx *ctx, int64_t *npairs);
/* Export the listed pair set as sorted (i,j) 1-based pairs, i<j [2*npairs]. */
int pgk_list_pairs(pgk_ctx *ctx, int32_t *pairs, int64_t capacity);

/* One calforce call (calforce.F90:8-1073): zero outputs, evaluate every term
 * whose MTS level equals mts_flag, add the net-charge terms on MTS_LONG, sum
 * the energies.  ifcalpreatom selects the virial-producing variants
 * (ewaldrp, calbondp, ...).  Positions must have been set.                  */
int pgk_calforce(pgk_ctx *ctx, int32_t mts_flag, const pgk_mts *mts,
                 int32_t ifcalpreatom, pgk_force_out *out);
/* One calforce_hf call on a heat-flux step (calforce_hf.F90:429-1451). */
int pgk_calforce_hf(pgk_ctx *ctx, int32_t mts_flag, const pgk_mts *mts,
                    const pgk_hf_desc *hf, pgk_force_out *out,
                    pgk_hf_out *hfo);

/* calheatf (calheatf.F90:8-1470) on the device, over the per-atom arrays and
 * the region description of the last pgk_calforce_hf call (which may skip
 * copying them out, see pgk_hf_out): per region r the kinetic and potential
 * transport sum_{centre_i in r} e_i v_i (molecule centre and centre-of-mass
 * velocity for MOLE atoms), the collision part sum_i sum_a W(a,n,i,r) v_i(a),
 * each divided by the region volume Lx Ly (z2 - z1), and their total.
 * Volume mode: z components only (the others are 0); bulk mode: all three.
 * vel [3*natom] and mass [natom] in units where m v^2 / 2 is in the energy
 * unit of the force path (the reference's non-dimensional atmvel, atmmass),
 * lab-frame (streaming velocities: pgk_set_atmvel_strm); mass NULL reuses
 * the last masses given.  hf [nregion*PGK_NHFJ*3].  The
 * surface mode keeps state across steps and is not offered.                 */
int pgk_calheatf(pgk_ctx *ctx, const double *vel, const double *mass, double *hf);

/* atmvel_strm (md_global.F90:443, assigned by ass_strmvel.F90:8-295): the
 * streaming velocity of each atom, [3*natom] in the units of vel, kept by
 * the context until changed.  pgk_calheatf and pgk_calforce_heatf take each
 * atom's kinetic energy as m (v - vs)^2 / 2 (calheatf.F90:368-369, 677-681);
 * the centre velocities they transport and the collision part use vel as
 * given (calheatf.F90:564, 644, 1023).  NULL clears it (zero).  The MD loop
 * uses the streaming slabs of pgk_md_langevin instead.                      */
int pgk_set_atmvel_strm(pgk_ctx *ctx, const double *vstrm);

/* calforce_hf followed by calheatf in one pass (no reference counterpart as
 * one call: the fused z-binned heat-flux reduction for hosts that only need
 * the region fluxes).  The heat-flux pass sums each interaction's share of
 * sum_a W(a,n,i,r) v_i(a) directly, fi.v_i - fj.v_j times the region weight
 * of calvirihf, so the (3,3,natom,nregion) arrays are never formed.  Forces,
 * energies and virials as pgk_calforce_hf; hf as pgk_calheatf.  Volume form
 * up to 1024 regions, or the bulk form; the order of the FP64 additions is
 * not fixed (hf agrees with pgk_calforce_hf + pgk_calheatf to rounding).    */
int pgk_calforce_heatf(pgk_ctx *ctx, int32_t mts_flag, const pgk_mts *mts,
                       const pgk_hf_desc *hf, const double *vel, const double *mass,
                       pgk_force_out *out, double *hf_out);

/* Device-resident variant used by GPU-resident drivers: same as pgk_calforce
 * but the force stays on the device (d_force, [3*natom] doubles, may be NULL)
 * and only the energy/virial records are copied to the host.                */
int pgk_calforce_device(pgk_ctx *ctx, int32_t mts_flag, const pgk_mts *mts,
                        int32_t ifcalpreatom, double *d_force, double *ene,
                        double *virt);

/* Queue one calforce on the context's stream and return without waiting
 * (no reference counterpart: the split of calforce for drivers that keep the
 * step on the device).  The force goes to d_force ([3*natom] doubles on the
 * device, NULL = the context's buffer); energies and virials are reduced on
 * the device and read with pgk_energies.  Calls queue in order.             */
int pgk_calforce_async(pgk_ctx *ctx, int32_t mts_flag, const pgk_mts *mts,
                       int32_t ifcalpreatom, double *d_force);
/* Wait for the context's work and return the energy record and virial
 * tensors of the last calforce call, with pgk_calforce's semantics (ene is
 * inout: energies of terms the call did not evaluate keep their values).    */
int pgk_energies(pgk_ctx *ctx, double *ene, double *virt);

/* Wait for all work queued by the context. */
int pgk_synchronize(pgk_ctx *ctx);

/* ---- device-resident NVE step (no reference counterpart in the force path:
 * the integrator of moldyn.F90:1465-1771 with the RATTLE of rattle.F90:6-272
 * (rattle_c) and rattle_mtk.F90:278-542 (rattlep_v), for drivers that keep the
 * whole step on the GPU) --------------------------------------------------- */
typedef struct pgk_md_desc {
  const double *mass;          /* [natom] amu                                  */
  int32_t ncons;               /* rigid bonds                                  */
  const int32_t *cons_ij;      /* [2*ncons] 1-based atom pairs                 */
  const double *cons_d;        /* [ncons] lengths (A)                          */
  int32_t ngroup;              /* coupled constraint groups (molecules)        */
  const int32_t *group_start;  /* [ngroup+1] 0-based offsets into the list    */
  double dt;                   /* time step (fs)                               */
  double tol;                  /* eps_rattle: rattle_c stops at |r^2-d^2|/d^2  *
                                * < tol, rattlep_v at (r.v)^2/r^2 v^2 < tol^2 */
} pgk_md_desc;
int pgk_md_setup(pgk_ctx *ctx, const pgk_md_desc *md);
/* Velocities [3*natom] (A/fs); the constraint velocity projection is applied. */
int pgk_md_set_velocities(pgk_ctx *ctx, const double *vel);
int pgk_md_get_state(pgk_ctx *ctx, double *pos, double *vel);
/* Run nstep velocity-Verlet steps from the current positions and velocities
 * (every force term on MTS_LONG; the Verlet list rebuilt every nstep_book
 * steps with rcut_book).  Every nsample steps (and at step 0) samples[3*k..]
 * receives kinetic, potential (pot_tot) and total energy (eref); samples
 * holds nstep/nsample + 1 rows.                                              */
int pgk_md_run(pgk_ctx *ctx, int32_t nstep, int32_t nstep_book, double rcut_book,
               int32_t nsample, double *samples);
/* Independent replicas stepped together (no reference counterpart: the
 * batched Green-Kubo ensemble).  Every ctxs[r] has had pgk_md_setup and
 * velocities; replica r uses rcut_book[r].  Up to 8 host threads share the
 * replicas (balanced by atom count; one thread when graph replay is on for
 * any replica); each queues step s of its replicas on
 * their own streams before step s+1 of any, so small replicas run side by
 * side.  samples: replica-major, (nstep/nsample + 1) rows of 3 per replica.
 * Each replica's trajectory is the one pgk_md_run gives.  The contexts must
 * not be used by other threads meanwhile.                                   */
int pgk_md_run_batch(pgk_ctx *const *ctxs, int32_t n, int32_t nstep, int32_t nstep_book,
                     const double *rcut_book, int32_t nsample, double *samples);
/* Replay steps that rebuild no list and sample nothing from a captured HIP
 * graph instead of launching their kernels one by one (no reference
 * counterpart: a switch for drivers and tests).  enable is a bit set: bit 0
 * for pgk_md_run, bit 1 for this context's replica in pgk_md_run_batch.
 * Both default off: on MI355X a graph launch cost more than the eager
 * launches it replaced, for the 100k box and for the 64-replica batch.      */
int pgk_md_use_graphs(pgk_ctx *ctx, int32_t enable);

/* Standard Ewald reciprocal sum with per-atom energy and virial: prepewk's
 * wave vectors (prepewk.F90:9-158, |n| <= kmax with the y/z ratios) and
 * ewaldkp_hf (ewaldk_hf.F90:8-304), the reciprocal part of calforce_hf in the
 * reference's _HF_BULK_EWK build, for small systems (natom x nwave work).
 * Uses the context's positions, charges, box and alpha.  As the reference:
 * adds into force [3*natom], pot_viri_coul, virt (pot_virit_coul, Fortran
 * 3x3), pot_atm (pot_elc_atm [natom], with the self term -alpha/sqrt(pi)
 * q_i^2) and viri_atm (virielct_atm(:,:,:,1) [9*natom]); sets pot_ewk.      */
int pgk_ewaldk_hf(pgk_ctx *ctx, int32_t kmax, double yratio, double zratio,
                  int32_t ifnetqcorrp, double *force, double *pot_ewk, double *pot_viri_coul,
                  double *virt, double *pot_atm, double *viri_atm);

/* Per-atom reciprocal (SPME) energy of the last call that evaluated the
 * reciprocal term at the current positions (no reference counterpart: the
 * SPME form of ewaldkp_hf's per-atom energy): e_i = q_i phi(r_i) / 2 with
 * phi the reciprocal potential interpolated from the same grid as the forces
 * (fft_pme.F90:872-937 with B-splines in place of their derivatives), so
 * sum_i e_i = pot_ewk.  e_atom [natom], eref; the self term is not included. */
int pgk_recip_energy_atom(pgk_ctx *ctx, double *e_atom);
/* pgk_recip_energy_atom without any device allocation (no reference
 * counterpart; for callers that cannot allocate: code under stream capture,
 * tight sampling loops, hosts that forbid allocation after set-up).  Same
 * arguments, result and errors; the per-atom energies are staged in a fixed
 * device array of PGK_NOALLOC_MAX_ATOMS = 1,048,576 doubles (8 MiB per
 * device, held until the process ends), so natom above that returns
 * PGK_EINVAL.  All calls on one device share that array: calls that overlap
 * on one device (other host threads or streams) race.                    */
int pgk_recip_energy_atom_noalloc(pgk_ctx *ctx, double *e_atom);
/* Per-atom reciprocal (SPME) virial tensors of the same evaluation (no
 * reference counterpart: the SPME form of ewaldkp_hf's virielct_atm):
 * W_i^ab = q_i phi_ab(r_i) / 2, with phi_ab the charge grid convolved with
 * scalar_sum's virial kernel eterm (delta_ab - vterm m_a m_b)
 * (fft_pme.F90:544-690; the sign of the Coulomb virial tensor and of
 * virielct_atm), one transform pair per tensor component.  viri_atm
 * [9*natom], Fortran (3,3,natom), eref; sum_i W_i is the reciprocal virial
 * tensor of the evaluation.  The net-charge correction (ifnetqcorrp) is not
 * distributed over the atoms.                                              */
int pgk_recip_virial_atom(pgk_ctx *ctx, double *viri_atm);

/* Heat-flux time series of pgk_md_run / pgk_md_run_batch (no reference
 * counterpart as a call: the device form of the heat-flux outputs that the
 * reference's moldyn_hf writes every heatfinterval steps, output_hf.F90:8-686,
 * from calforce_hf + calheatf).  At step 0 and every `interval` steps the
 * fused pass of pgk_calforce_heatf runs at the step's end state (positions
 * and velocities after the second half kick and RATTLE), every atom its own
 * centre (hftyp NULL or all ATOM), with the MD units: velocities in A/fs,
 * m v^2 / 2 in eref, so the flux is in eref A fs^-1 / A^3.  Rows
 * [nregion*PGK_NHFJ*3] stay on the device until pgk_md_heatf_series.  hf
 * NULL or interval <= 0 turns it off.                                       */
int pgk_md_heatf(pgk_ctx *ctx, const pgk_hf_desc *hf, int32_t interval);
/* Rows of the last run (or batch): nrows receives their number; up to
 * max_rows of them are copied to out.                                       */
int pgk_md_heatf_series(pgk_ctx *ctx, double *out, int32_t max_rows, int32_t *nrows);

/* RESPA multiple time stepping for pgk_md_run / pgk_md_run_batch
 * (moldyn.F90:1465-1771): nstep_short inner steps of dt/nstep_short per
 * outer step; the long-level forces' half kick opens the first inner step and
 * closes the last, they are evaluated once per outer step after the last
 * drift; the short-level forces every inner step; rattle_c at the inner step,
 * rattlep_v after every inner step.  mts gives each term's level
 * (PGK_MTS_LONG, PGK_MTS_SHORT or 0; NULL = all long); nstep_short 1 turns
 * it off.  The sampled potential energy is the short call's plus the long
 * call's.  Not combined with the Nose-Hoover chain.                          */
int pgk_md_respa(pgk_ctx *ctx, int32_t nstep_short, const pgk_mts *mts);

/* Nose-Hoover chain thermostat for pgk_md_run / pgk_md_run_batch: the NVT
 * step of moldyn_nhc.F90:1238-1690 with every term on the long step, i.e.
 * nhcint (nhcint.F90:8-140), the thermostatted velocity update
 * v = v exp(-dt vxi1/2) + F/m bbv (moldyn_nhc.F90:1285-1302, 1589-1590),
 * drift, rattle_c, forces, the same velocity update, rattlep_v, nhcint.
 * Q1 = nfree kT / tfreq^2, Qi = kT / tfreq^2 (prepmd.F90:515-518); the
 * Yoshida-Suzuki weights of prepmd.F90:534-548.  With the chain on, the
 * total-energy sample is the conserved quantity of calkin.F90:189-201
 * (KE + PE + sum Q vxi^2/2 + nfree kT xi1 + kT sum_{i>=2} xi_i).          */
typedef struct pgk_nhc_desc {
  int32_t mchain;              /* chain length, 2..10 (maxnhc)                */
  double temperature;          /* text_c (K)                                  */
  double tfreq;                /* thermostat frequency (1/fs)                 */
  double nfree;                /* degfree_all: degrees of freedom             */
  int32_t nyosh;               /* Yoshida-Suzuki order: 1, 3 or 5             */
  int32_t next;                /* nhcint sub-steps (next >= 1)                */
  const double *xlogs;         /* [mchain] start xi (NULL = 0)                */
  const double *vlogs;         /* [mchain] start vxi (1/fs, NULL = 0)         */
} pgk_nhc_desc;
/* nhc NULL switches the thermostat off (plain NVE).                          */
int pgk_md_nhc(pgk_ctx *ctx, const pgk_nhc_desc *nhc);

/* Region Langevin thermostat for pgk_md_run / pgk_md_run_batch (the
 * ifreglange option of the reference's NVE and RESPA drivers,
 * moldyn.F90 / moldyn_hf.F90:1667-1678, 2348-2372).  At the end of step 0
 * and of every step the atoms are assigned to the first region whose box
 * [x1,x2) x [y1,y2) x [z1,z2) holds them (langevin_regist.F90:8-364): a
 * molecule flagged in mol_com by its centre of mass, moving as a whole,
 * other atoms one by one; coordinates are wrapped once into [0, L) and a
 * box may reach past the cell by one period.  At the start of each step
 * every registered atom draws a random acceleration by the Box-Muller method
 * with variance 2 damp kT / (m dt) per component and gets
 * v = (v + dt/2 a) exp(-dt damp/2); at the end of the step, after RATTLE,
 * v = v exp(-dt damp/2) + dt/2 a with the same a (zero for atoms that joined
 * a region during the step) (langevin_region.F90:8-203).  The uniform
 * numbers come from SplitMix64 at counter positions
 * seed + (4 (draw natom + i) + k + 1) 0x9E3779B97F4A7C15, mapped to
 * ((z >> 11) + 1) 2^-53 in (0, 1], in place of the reference's SFMT stream;
 * draw counts the steps run since pgk_md_langevin.  With the Nose-Hoover
 * chain on as well, the chain acts on the atoms outside every region (its
 * kinetic energy and its half kicks) and the regions keep their own
 * thermostat.  lt NULL switches the thermostat and the streaming slabs off. */
#define PGK_LT_MAXREG 16
typedef struct pgk_langevin_desc {
  int32_t nregion;             /* 0..PGK_LT_MAXREG (nlangeregion); 0: no      *
                                * thermostat, streaming slabs only (needs     *
                                * nstream >= 1; the region arrays may be NULL)*/
  const double *x1, *x2;       /* [nregion] ltxpos1 / ltxpos2 (A)             */
  const double *y1, *y2;       /* [nregion] ltypos1 / ltypos2 (A)             */
  const double *z1, *z2;       /* [nregion] ltzpos1 / ltzpos2 (A)             */
  const double *temperature;   /* [nregion] r_ltemp (K)                       */
  const double *damp;          /* [nregion] r_ltdamp (1/fs)                   */
  const int32_t *mol_com;      /* [nmol] 1: molecule registered by its centre *
                                * of mass (water; polymers unless iftcratom), *
                                * 0: atom by atom; NULL: every molecule by    *
                                * its centre of mass.  Molecules are molid's. */
  uint64_t seed;
  /* streaming velocity (ass_strmvel.F90:8-295, ifstrmvel): nstream slabs
   * [z1,z2) in the wrapped z of the same registration units; an atom whose
   * unit lies in slab s has streaming velocity strm_vel[3s..3s+2] (A/fs),
   * assigned after every drift at the new positions (moldyn_nhc.F90:1412).
   * Everything that measures or controls the temperature sees v minus it:
   * the region thermostat (langevin_region.F90:81-82, 190), the chain's
   * kinetic energy (nhcint.F90:53-54) and half kicks, which scale v - vs
   * and add vs back (moldyn_nhc.F90:1280-1310, 1585-1598), the kinetic
   * energy of the sampled rows (calkin.F90:98) and the atoms' kinetic
   * energies in the MD loop's heat flux (calheatf.F90:368-369, whose
   * transport velocities and collision part stay lab-frame).
   * nstream 0 (or NULL arrays): none.                                       */
  int32_t nstream;             /* 0..PGK_LT_MAXREG                            */
  const double *strm_z1, *strm_z2;   /* [nstream] strmzpos1 / strmzpos2 (A)   */
  const double *strm_vel;      /* [3*nstream] strmvel (A/fs)                  */
} pgk_langevin_desc;
int pgk_md_langevin(pgk_ctx *ctx, const pgk_langevin_desc *lt);
/* Thermostat record: det_ene_kin[nregion] the kinetic energy (eref) the
 * thermostat gave each region over the last run (langevin_region.F90's
 * det_ene_kin summed over the run's steps); natom_lt[nregion] the atoms in
 * each region at the last registration.  Either pointer may be NULL.       */
int pgk_md_langevin_state(pgk_ctx *ctx, double *det_ene_kin, int32_t *natom_lt);
/* Current chain state: xlogs[mchain], vlogs[mchain] (1/fs); NULL skips.      */
int pgk_md_nhc_state(pgk_ctx *ctx, double *xlogs, double *vlogs);

/* ---- on-disk formats (host only; no device needed) ------------------------
 * The state file of wrsta / rdstarec (wrsta.F90:8-180, rdstarec.F90:8-393):
 * counts, box (m), step, one line per atom (x y z in m, vx vy vz in m/s,
 * atom type) in three sections (poly, water, monatomic molecules), the
 * Nose-Hoover chain (xlogs, vlogs in 1/s), the Andersen barostat (vlogv in
 * 1/s, pint in Pa) and its anisotropic form (vboxg(3) in 1/s, pintt(3,3) in
 * Pa).  Written with the reference's edit descriptors, so a file read and
 * written back is byte-identical.                                            */
typedef struct pgk_sta_header {
  int32_t npoly, nwater, nmatom, natom, mchain;
  int32_t nat_poly, nat_water, nat_matom;   /* atom lines in each section    */
  double box[3];                            /* xcel, ycel, zcel (m)          */
  int64_t nstep;
} pgk_sta_header;
int pgk_sta_read_header(const char *path, pgk_sta_header *h);
/* pos, vel [3*natom]; types [4*natom] (each up to 3 characters, NUL-padded);
 * nhc [2*mchain] (xlogs, vlogs pairs); andersen [2]; aniso [12].            */
int pgk_sta_read(const char *path, const pgk_sta_header *h, double *pos, double *vel,
                 char *types, double *nhc, double *andersen, double *aniso);
int pgk_sta_write(const char *path, const pgk_sta_header *h, const double *pos,
                  const double *vel, const char *types, const double *nhc,
                  const double *andersen, const double *aniso);
/* One data row of out_ene.dat as outene writes it (output_data.F90:9-342):
 * (I10, nvals(1X,E15.8)), or E23.16 with high != 0 (-D_OUTENE_HIGH), with a
 * newline, into buf (cap bytes).                                             */
int pgk_out_ene_row(int32_t step, const double *vals, int32_t nvals, int32_t high, char *buf,
                    int32_t cap);
/* out_pre.dat as outpre writes it (output_data.F90:549-818), for a build
 * without custom non-bonded terms (ncstmnbex = 0): the header line of its
 * first call, (A9,1X,81(1X,A23)) with the field names right-justified, and
 * one data row (I10, nvals(1X,E23.8)), E23.16 with high != 0
 * (-D_OUTENE_HIGH).  The caller supplies the 81 values in the file's order,
 * already in Pa / J.                                                        */
int pgk_out_pre_header(char *buf, int32_t cap);
int pgk_out_pre_row(int32_t step, const double *vals, int32_t nvals, int32_t high, char *buf,
                    int32_t cap);
/* out_htf.dat as outhtf writes it (output_hf.F90:8-686), for a build
 * without custom non-bonded terms: 36 fields per region (total, kinetic,
 * the 13 potential and 13 virial classes, the 8 argon-pair splits), each
 * with x, y, z of every region when all_dir != 0 (_HF_ALL_DIR / _HF_BULK
 * builds) or z only.  The header names are name_<c>_<region> written
 * left-justified in A23; a row takes hf as pgk_calheatf returns it
 * ([nregion][PGK_NHFJ][3]) times jref, with the classes this library does
 * not evaluate (Morse, SH, RFH, DOU, custom, argon splits) as zero,
 * (I10, n(1X,E23.8)) or E23.16 with high != 0 (-D_OUTHF_HIGH).  As the
 * reference, a surface-mode driver writes no row at its first step.       */
int pgk_out_htf_header(int32_t nregion, int32_t all_dir, char *buf, int32_t cap);
int pgk_out_htf_row(int32_t step, const double *hf, int32_t nregion, int32_t all_dir,
                    int32_t high, double jref, char *buf, int32_t cap);

/* The .ini control script (rdscript.F90 with rdfree.F90's word splitting):
 * each line's words split at spaces, commas and tabs, the first word the key,
 * reading stops at a line starting with END.  Keys may repeat (occurrence
 * occ counts from 0); value idx 0 is the word after the key.  Reals accept
 * Fortran d/D exponents, logicals [.]T... / [.]F... as list-directed input. */
typedef struct pgk_ini pgk_ini;
int pgk_ini_read(const char *path, pgk_ini **ini);
void pgk_ini_free(pgk_ini *ini);
int32_t pgk_ini_count(const pgk_ini *ini, const char *key);
int32_t pgk_ini_nvalues(const pgk_ini *ini, const char *key, int32_t occ);  /* -1: absent */
int pgk_ini_get_str(const pgk_ini *ini, const char *key, int32_t occ, int32_t idx, char *buf,
                    int32_t cap);
int pgk_ini_get_double(const pgk_ini *ini, const char *key, int32_t occ, int32_t idx,
                       double *v);
int pgk_ini_get_int(const pgk_ini *ini, const char *key, int32_t occ, int32_t idx, int64_t *v);
int pgk_ini_get_bool(const pgk_ini *ini, const char *key, int32_t occ, int32_t idx,
                     int32_t *v);

/* Stage timing (no reference counterpart: instrumentation for measurement).
 * enable != 0 turns HIP-event timing of each stage on (and clears the
 * totals); ms[PGK_NPROF_STAGES] and calls[PGK_NPROF_STAGES] receive the
 * accumulated device time and launch count of each stage.                    */
enum {
  PGK_PROF_PAIRS = 0, PGK_PROF_EXCL, PGK_PROF_SPREAD, PGK_PROF_FFT,
  PGK_PROF_GATHER, PGK_PROF_BONDED, PGK_PROF_LIST, PGK_PROF_OTHER,
  PGK_NPROF_STAGES
};
int pgk_profile(pgk_ctx *ctx, int32_t enable, double *ms, int64_t *calls);
/* Stages that record events while timing is on: bit k for stage k (default
 * all).  Timing fewer stages keeps the host's enqueue rate up.               */
int pgk_profile_stages(pgk_ctx *ctx, uint32_t mask);
/* Launch-shape choices (no reference counterpart: tuning).  None changes a
 * result beyond the rounding of the mixed and FP32 forms' FP32 partial sums
 * (the FP64 form's order of atomic additions is not fixed anyway).
 *   PGK_TUNE_PAIR_CHUNK_BELOW: systems with fewer 8-atom clusters than this
 *     run the pair kernel in its chunked form (each wave takes 64 list entries
 *     of one cluster instead of the cluster's whole list), which fills the
 *     device for small systems; 0 never, a huge value always.  Default
 *     PGK_TUNE_DEFAULT_CHUNK_BELOW, the crossover thrid_amd/autotune.py
 *     measured (thrid_amd/tuning.json).
 *   PGK_TUNE_PME_OVERLAP: 1 (default) runs the reciprocal chain on a second
 *     stream beside the pair kernel when both are due, 0 one after the other.
 *   PGK_TUNE_PAIR_CAP_BLOCKS: pair blocks per CU while the chain runs beside
 *     them (default 2; 0 no cap); the chunked form keeps 3 unless 0.        */
enum { PGK_TUNE_PAIR_CHUNK_BELOW = 1, PGK_TUNE_PME_OVERLAP = 2, PGK_TUNE_PAIR_CAP_BLOCKS = 3 };
#define PGK_TUNE_DEFAULT_CHUNK_BELOW 4352
int pgk_tune(pgk_ctx *ctx, int32_t knob, int64_t value);
/* Shape of the current list (no reference counterpart: instrumentation):
 * info[0] 8-atom clusters, info[1] list entries (an i-cluster against half a
 * j-cluster), info[2] listed atom pairs, info[3] pair slots per entry.       */
int pgk_list_info(pgk_ctx *ctx, int64_t *info);

#ifdef __cplusplus
}
#endif
#endif /* PGK_FORCE_H */
