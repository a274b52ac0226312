!-----------------------------------------------------------------------
! pgk_force_mod.F90 - Fortran binding of libpgkforce.so (iso_c_binding).
!
! The module a Fortran MD host links to call the MI355X force path in place
! of calforce / calforce_hf (calforce.F90:8-1073, calforce_hf.F90:8-1451).
! Types mirror include/pgk_force.h field for field (bind(C)); functions are
! the C entry points with the same names.  Arrays are passed with c_loc()
! of contiguous, TARGET host arrays; nothing here owns device memory.
!
! Typical use (see example_calforce.F90):
!   desc%natom = natom; desc%charge = c_loc(atmchrg); ...
!   call pgk_check(pgk_create(desc, PGK_PREC_FP64, ctx), 'pgk_create')
!   call pgk_check(pgk_set_box(ctx, box), 'pgk_set_box')
!   call pgk_check(pgk_set_positions(ctx, atmcor), 'pgk_set_positions')
!   call pgk_check(pgk_build_list(ctx, rcut_book), 'pgk_build_list')
!   call pgk_check(pgk_calforce(ctx, mts_flag, mts, 1, out), 'pgk_calforce')
!-----------------------------------------------------------------------
module pgk_force_mod
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: PGK_ABI = 1   ! PGK_ABI_VERSION of the header
  integer(c_int), parameter :: PGK_OK = 0, PGK_EINVAL = -1, PGK_EDEVICE = -2, &
       &                        PGK_ENOMEM = -3, PGK_ESTATE = -4, PGK_EOVERFLOW = -5
  integer(c_int32_t), parameter :: PGK_PREC_FP64 = 0, PGK_PREC_MIXED = 1, PGK_PREC_FP32 = 2
  integer(c_int32_t), parameter :: PGK_MTS_LONG = 1, PGK_MTS_MED = 2, PGK_MTS_SHORT = 3
  integer(c_int32_t), parameter :: PGK_HF_VOLUME = 0, PGK_HF_SURFACE = 1, PGK_HF_BULK = 2
  integer(c_int32_t), parameter :: PGK_HFTYP_ATOM = 1, PGK_HFTYP_MOLE = 2
  ! pgk_tune knobs and the compiled default of the pair-form crossover
  integer(c_int32_t), parameter :: PGK_TUNE_PAIR_CHUNK_BELOW = 1, PGK_TUNE_PME_OVERLAP = 2, &
       &                           PGK_TUNE_PAIR_CAP_BLOCKS = 3
  integer(c_int64_t), parameter :: PGK_TUNE_DEFAULT_CHUNK_BELOW = 4352
  integer, parameter :: PGK_NENE = 19, PGK_NTENS = 6, PGK_NHF = 8, PGK_NHFJ = 18
  ! energy record (ene(1:PGK_NENE), 1-based here; C index + 1)
  integer, parameter :: PGK_E_TOT = 1, PGK_E_NONBON = 2, PGK_E_VDW = 3, PGK_E_ELC = 4, &
       &                PGK_E_EWK = 5, PGK_E_EWC = 6, PGK_E_EWNQ = 7, PGK_E_VDW14 = 8, &
       &                PGK_E_ELC14 = 9, PGK_E_BOND = 10, PGK_E_ANGL = 11, PGK_E_TORS = 12, &
       &                PGK_E_TORSRB = 13, PGK_V_COUL = 14, PGK_V_LJ = 15, PGK_V_BOND = 16, &
       &                PGK_V_ANGL = 17, PGK_V_TORS = 18, PGK_V_14 = 19

  type, bind(C) :: pgk_system_desc
     integer(c_int32_t) :: natom = 0, ntype = 0
     type(c_ptr) :: charge = c_null_ptr, atype = c_null_ptr, molid = c_null_ptr
     type(c_ptr) :: lj_eps = c_null_ptr, lj_rmin = c_null_ptr, inttyp = c_null_ptr
     integer(c_int32_t) :: nexcl = 0
     type(c_ptr) :: excl_index = c_null_ptr, excl_list = c_null_ptr
     integer(c_int32_t) :: nbond = 0, nbondtyp = 0
     type(c_ptr) :: bond_ij = c_null_ptr, bond_type = c_null_ptr
     type(c_ptr) :: bond_k = c_null_ptr, bond_r0 = c_null_ptr
     integer(c_int32_t) :: nangl = 0, nangltyp = 0
     type(c_ptr) :: angl_ijk = c_null_ptr, angl_type = c_null_ptr
     type(c_ptr) :: angl_k = c_null_ptr, angl_theta0 = c_null_ptr
     integer(c_int32_t) :: ntors = 0, ntorstyp = 0
     type(c_ptr) :: tors_ijkl = c_null_ptr, tors_type = c_null_ptr, tors_ring14 = c_null_ptr
     type(c_ptr) :: tors_divfac = c_null_ptr, tors_barhig = c_null_ptr
     type(c_ptr) :: tors_phsang = c_null_ptr, tors_period = c_null_ptr
     type(c_ptr) :: tors_divfac_vdw = c_null_ptr, tors_divfac_elc = c_null_ptr
     type(c_ptr) :: tors_if14 = c_null_ptr
     integer(c_int32_t) :: ntorsrb = 0, ntorsrbtyp = 0
     type(c_ptr) :: torsrb_ijkl = c_null_ptr, torsrb_type = c_null_ptr
     type(c_ptr) :: torsrb_ring14 = c_null_ptr, torsrb_c = c_null_ptr
     type(c_ptr) :: torsrb_divfac_vdw = c_null_ptr, torsrb_divfac_elc = c_null_ptr
     type(c_ptr) :: torsrb_if14 = c_null_ptr
     real(c_double) :: alpha = 0.0d0, rrcut = 0.0d0, rcut = 0.0d0
     integer(c_int32_t) :: nspltbl = 0, nfft1 = 0, nfft2 = 0, nfft3 = 0
     integer(c_int32_t) :: pme_order = 4, ifnetqcorrp = 0
     type(c_ptr) :: box = c_null_ptr
  end type pgk_system_desc

  type, bind(C) :: pgk_mts
     integer(c_int32_t) :: bond = PGK_MTS_LONG, angl = PGK_MTS_LONG, tors = PGK_MTS_LONG
     integer(c_int32_t) :: torsrb = PGK_MTS_LONG, vdw = PGK_MTS_LONG, ewk = PGK_MTS_LONG
     integer(c_int32_t) :: nonb14 = PGK_MTS_LONG
  end type pgk_mts

  type, bind(C) :: pgk_force_out
     type(c_ptr) :: force = c_null_ptr, ene = c_null_ptr, virt = c_null_ptr
     type(c_ptr) :: for_viri_coul = c_null_ptr, for_viri_lj = c_null_ptr
  end type pgk_force_out

  type, bind(C) :: pgk_hf_desc
     integer(c_int32_t) :: mode = PGK_HF_VOLUME, nregion = 0
     type(c_ptr) :: z1 = c_null_ptr, z2 = c_null_ptr, hftyp = c_null_ptr
     integer(c_int32_t) :: nmol = 0
     type(c_ptr) :: molecom = c_null_ptr
  end type pgk_hf_desc

  type, bind(C) :: pgk_hf_out
     type(c_ptr) :: pot_atm = c_null_ptr, viri_atm = c_null_ptr
  end type pgk_hf_out

  interface
     function pgk_abi_version() bind(C, name='pgk_abi_version') result(v)
       import :: c_int32_t
       integer(c_int32_t) :: v
     end function pgk_abi_version

     function pgk_last_error() bind(C, name='pgk_last_error') result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function pgk_last_error

     function pgk_create(desc, precision, ctx) bind(C, name='pgk_create') result(rc)
       import :: pgk_system_desc, c_int32_t, c_ptr, c_int
       type(pgk_system_desc), intent(in) :: desc
       integer(c_int32_t), value :: precision
       type(c_ptr), intent(out) :: ctx
       integer(c_int) :: rc
     end function pgk_create

     function pgk_destroy(ctx) bind(C, name='pgk_destroy') result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function pgk_destroy

     function pgk_set_box(ctx, box) bind(C, name='pgk_set_box') result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: box(3)
       integer(c_int) :: rc
     end function pgk_set_box

     ! pos: Fortran atmcor(3, natom)
     function pgk_set_positions(ctx, pos) bind(C, name='pgk_set_positions') result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: pos(*)
       integer(c_int) :: rc
     end function pgk_set_positions

     function pgk_build_list(ctx, rcut_book) bind(C, name='pgk_build_list') result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: rcut_book
       integer(c_int) :: rc
     end function pgk_build_list

     function pgk_list_npairs(ctx, npairs) bind(C, name='pgk_list_npairs') result(rc)
       import :: c_ptr, c_int, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: npairs
       integer(c_int) :: rc
     end function pgk_list_npairs

     function pgk_calforce(ctx, mts_flag, mts, ifcalpreatom, out) &
          & bind(C, name='pgk_calforce') result(rc)
       import :: c_ptr, c_int, c_int32_t, pgk_mts, pgk_force_out
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: mts_flag
       type(pgk_mts), intent(in) :: mts
       integer(c_int32_t), value :: ifcalpreatom
       type(pgk_force_out), intent(in) :: out
       integer(c_int) :: rc
     end function pgk_calforce

     function pgk_calforce_hf(ctx, mts_flag, mts, hf, out, hfo) &
          & bind(C, name='pgk_calforce_hf') result(rc)
       import :: c_ptr, c_int, c_int32_t, pgk_mts, pgk_hf_desc, pgk_force_out, pgk_hf_out
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: mts_flag
       type(pgk_mts), intent(in) :: mts
       type(pgk_hf_desc), intent(in) :: hf
       type(pgk_force_out), intent(in) :: out
       type(pgk_hf_out), intent(in) :: hfo
       integer(c_int) :: rc
     end function pgk_calforce_hf

     ! calheatf over the last pgk_calforce_hf call's arrays (which that call
     ! may leave on the device: hfo%pot_atm = hfo%viri_atm = c_null_ptr);
     ! hf(3, PGK_NHFJ, nhfregion) = hfkin, hfpot*, hfviri*, hftot
     function pgk_calheatf(ctx, vel, mass, hf) bind(C, name='pgk_calheatf') result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: vel(*), mass(*)
       real(c_double), intent(out) :: hf(*)
       integer(c_int) :: rc
     end function pgk_calheatf

     ! atmvel_strm (md_global.F90:443) for pgk_calheatf / pgk_calforce_heatf:
     ! vstrm = c_loc(atmvel_strm), or c_null_ptr to clear it
     function pgk_set_atmvel_strm(ctx, vstrm) bind(C, name='pgk_set_atmvel_strm') result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, vstrm
       integer(c_int) :: rc
     end function pgk_set_atmvel_strm

     ! launch-shape choices of one context (no reference counterpart)
     function pgk_tune(ctx, knob, value) bind(C, name='pgk_tune') result(rc)
       import :: c_ptr, c_int, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: knob
       integer(c_int64_t), value :: value
       integer(c_int) :: rc
     end function pgk_tune

     ! calforce_hf and calheatf in one pass, without the per-atom arrays
     function pgk_calforce_heatf(ctx, mts_flag, mts, hf, vel, mass, out, hfj) &
          & bind(C, name='pgk_calforce_heatf') result(rc)
       import :: c_ptr, c_int, c_int32_t, c_double, pgk_mts, pgk_hf_desc, pgk_force_out
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: mts_flag
       type(pgk_mts), intent(in) :: mts
       type(pgk_hf_desc), intent(in) :: hf
       real(c_double), intent(in) :: vel(*), mass(*)
       type(pgk_force_out), intent(in) :: out
       real(c_double), intent(out) :: hfj(*)
       integer(c_int) :: rc
     end function pgk_calforce_heatf

     ! ewaldkp_hf with prepewk's wave vectors (the _HF_BULK_EWK build)
     function pgk_ewaldk_hf(ctx, kmax, yratio, zratio, ifnetqcorrp, force, pot_ewk, &
          &                 pot_viri_coul, virt, pot_atm, viri_atm) &
          & bind(C, name='pgk_ewaldk_hf') result(rc)
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: kmax, ifnetqcorrp
       real(c_double), value :: yratio, zratio
       real(c_double), intent(inout) :: force(*), pot_viri_coul, virt(*), pot_atm(*), viri_atm(*)
       real(c_double), intent(out) :: pot_ewk
       integer(c_int) :: rc
     end function pgk_ewaldk_hf

     ! per-atom SPME reciprocal energies (natom) and virial tensors
     ! (3,3,natom, virielct_atm's layout) of the last reciprocal evaluation
     function pgk_recip_energy_atom(ctx, e_atom) &
          & bind(C, name='pgk_recip_energy_atom') result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: e_atom(*)
       integer(c_int) :: rc
     end function pgk_recip_energy_atom

     function pgk_recip_virial_atom(ctx, viri_atm) &
          & bind(C, name='pgk_recip_virial_atom') result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: viri_atm(*)
       integer(c_int) :: rc
     end function pgk_recip_virial_atom

     ! one line of out_htf.dat (outhtf) from hf(3, PGK_NHFJ, nhfregion) times
     ! jref, and its header; buf receives a NUL-terminated line
     function pgk_out_htf_header(nregion, all_dir, buf, cap) &
          & bind(C, name='pgk_out_htf_header') result(rc)
       import :: c_int, c_int32_t, c_char
       integer(c_int32_t), value :: nregion, all_dir, cap
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int) :: rc
     end function pgk_out_htf_header

     function pgk_out_htf_row(step, hf, nregion, all_dir, high, jref, buf, cap) &
          & bind(C, name='pgk_out_htf_row') result(rc)
       import :: c_int, c_int32_t, c_double, c_char
       integer(c_int32_t), value :: step, nregion, all_dir, high, cap
       real(c_double), intent(in) :: hf(*)
       real(c_double), value :: jref
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int) :: rc
     end function pgk_out_htf_row

     function pgk_calforce_async(ctx, mts_flag, mts, ifcalpreatom, d_force) &
          & bind(C, name='pgk_calforce_async') result(rc)
       import :: c_ptr, c_int, c_int32_t, pgk_mts
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: mts_flag
       type(pgk_mts), intent(in) :: mts
       integer(c_int32_t), value :: ifcalpreatom
       type(c_ptr), value :: d_force
       integer(c_int) :: rc
     end function pgk_calforce_async

     function pgk_energies(ctx, ene, virt) bind(C, name='pgk_energies') result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: ene(*)
       real(c_double), intent(out) :: virt(*)
       integer(c_int) :: rc
     end function pgk_energies

     function pgk_synchronize(ctx) bind(C, name='pgk_synchronize') result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function pgk_synchronize
  end interface

contains

  ! The last error of the calling thread as a Fortran string.
  function pgk_error_string() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: n
    p = pgk_last_error()
    if (.not. c_associated(p)) then
       msg = ''
       return
    end if
    call c_f_pointer(p, cs, [4096])
    n = 0
    do while (n < 4096)
       if (cs(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(cs(1:n), msg)
  end function pgk_error_string

  ! The reference's error convention (write(6,*) ...; stop) for a failed call.
  subroutine pgk_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= PGK_OK) then
       write(6, '(a, a, i0, a, a)') where, ' failed (', rc, '): ', pgk_error_string()
       stop 1
    end if
  end subroutine pgk_check

end module pgk_force_mod
