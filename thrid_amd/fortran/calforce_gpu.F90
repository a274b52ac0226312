!-----------------------------------------------------------------------
! calforce_gpu.F90 - drop-in replacements of peachgk_md's calforce,
! calforce_hf, mklist2a and mklist_cell that run on the MI355X library
! (libpgkforce.so) through pgk_force_mod.
!
! The four subroutines keep the reference's names and dummy argument lists
! (interface_mdtech.F90:154-181, 358-608, 609-951), so a host built with
! them in place of calforce.o, calforce_hf.o, mklist.o and mklist_cell.o calls
! the GPU path with its callers untouched (moldyn*.F90, enemin_sd.F90).  They
! read the system from the host's module md_global as the reference routines
! do: atmcor, atmchrg, atmindex, irmolept_list, vdw_welij / vdw_radij /
! inter_inttyp, the exclusion CSR of mkexcl and the bonded lists and
! parameters.
!
! The library context is created at the first call from those globals and
! kept for the run (recreated if natom changes).  mklist2a / mklist_cell build
! the GPU neighbour list at the current positions (the reference's pair set);
! calforce / calforce_hf upload the current positions and box and evaluate.
! The arithmetic form comes from the environment variable PGK_PRECISION
! (fp64 (default), mixed or fp32), read once.
!
! The library evaluates the terms of the reference's default build; a host
! whose input switches on a term it does not evaluate (Urey-Bradley angles,
! improper torsions, Morse, SH, RFH, DOU, custom non-bonded, position
! restraints, bias potentials, standard Ewald) stops with a message instead of
! silently leaving it out.  Errors stop as the reference does (write; stop).
!-----------------------------------------------------------------------
module pgk_shim_state
  use, intrinsic :: iso_c_binding
  use pgk_force_mod
  implicit none
  type(c_ptr), save :: shim_ctx = c_null_ptr
  integer, save :: shim_natom = -1
  real(c_double), save :: shim_box(3) = 0.0d0
  real(c_double), save :: shim_pending_book = -1.0d0   ! list requested before the context
  logical, save :: shim_have_list = .false.
  ! host copies the descriptor points into (kept for the context's life)
  real(c_double), allocatable, target, save :: s_chrg(:), s_eps(:), s_rmin(:)
  integer(c_int32_t), allocatable, target, save :: s_typ(:), s_mol(:), s_ityp(:)
  integer(c_int32_t), allocatable, target, save :: s_eidx(:), s_elst(:)
  integer(c_int32_t), allocatable, target, save :: s_bij(:), s_btyp(:)
  real(c_double), allocatable, target, save :: s_bk(:), s_br0(:)
  integer(c_int32_t), allocatable, target, save :: s_aijk(:), s_atyp(:)
  real(c_double), allocatable, target, save :: s_ak(:), s_at0(:)
  integer(c_int32_t), allocatable, target, save :: s_tijkl(:), s_ttyp(:), s_tdiv(:), s_t14(:)
  real(c_double), allocatable, target, save :: s_tring(:), s_tbar(:), s_tphs(:), s_tper(:)
  real(c_double), allocatable, target, save :: s_tdv(:), s_tde(:)
  integer(c_int32_t), allocatable, target, save :: s_rijkl(:), s_rtyp(:), s_r14(:)
  real(c_double), allocatable, target, save :: s_rring(:), s_rc(:), s_rdv(:), s_rde(:)
  real(c_double), target, save :: s_boxd(3) = 0.0d0
contains

  subroutine shim_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    call pgk_check(rc, 'calforce_gpu: ' // what)
  end subroutine shim_check

  integer(c_int32_t) function shim_precision() result(p)
    character(len=16) :: v
    integer :: ln, st
    p = PGK_PREC_FP64
    call get_environment_variable('PGK_PRECISION', v, ln, st)
    if (st /= 0) return
    select case (v(1:ln))
    case ('mixed', 'MIXED')
       p = PGK_PREC_MIXED
    case ('fp32', 'FP32')
       p = PGK_PREC_FP32
    case ('fp64', 'FP64', '')
       p = PGK_PREC_FP64
    case default
       write(6,*) 'calforce_gpu: PGK_PRECISION must be fp64, mixed or fp32'
       stop
    end select
  end function shim_precision

  ! The context from md_global (first force call, or natom changed); alpha,
  ! the cutoffs and the SPME grid come from calforce's arguments.
  subroutine shim_ensure(xcel, ycel, zcel, alpha, rrcut, rcut, nfft1, nfft2, nfft3, &
       &                 pme_order)
    use md_global
    real(8), intent(in) :: xcel, ycel, zcel, alpha, rrcut, rcut
    integer, intent(in) :: nfft1, nfft2, nfft3, pme_order
    type(pgk_system_desc) :: d
    integer :: nt, i, k
    if (c_associated(shim_ctx) .and. shim_natom == natom) then
       call shim_box_update(xcel, ycel, zcel)
       return
    end if
    if (c_associated(shim_ctx)) call shim_check(pgk_destroy(shim_ctx), 'pgk_destroy')
    shim_ctx = c_null_ptr
    if (nanglub > 0 .or. ntorsim > 0) then
       write(6,*) 'calforce_gpu: Urey-Bradley angles and improper torsions are not ', &
            &     'evaluated by the GPU path'
       stop
    end if
    nt = maxval(atmindex(1:natom))
    call shim_free()
    allocate(s_chrg(natom), s_typ(natom), s_mol(natom))
    s_chrg(1:natom) = atmchrg(1:natom)
    s_typ(1:natom) = atmindex(1:natom)
    s_mol(1:natom) = irmolept_list(1:natom)
    allocate(s_eps(nt*nt), s_rmin(nt*nt), s_ityp(nt*nt))
    do k = 1, nt
       do i = 1, nt
          s_eps(i + nt*(k-1)) = vdw_welij(i, k)
          s_rmin(i + nt*(k-1)) = vdw_radij(i, k)
          s_ityp(i + nt*(k-1)) = inter_inttyp(i, k)
       end do
    end do
    allocate(s_eidx(natom+1), s_elst(max(nlistexcl, 1)))
    s_eidx(1:natom+1) = index_excl(1:natom+1)
    s_elst = 0
    s_elst(1:nlistexcl) = list_excl(1:nlistexcl)
    allocate(s_bij(max(2*nbond, 1)), s_btyp(max(nbond, 1)))
    allocate(s_bk(max(nbondtyp, 1)), s_br0(max(nbondtyp, 1)))
    s_bij = 0; s_btyp = 0; s_bk = 0.0d0; s_br0 = 0.0d0
    do i = 1, nbond
       s_bij(2*i-1) = ibond(i)
       s_bij(2*i) = jbond(i)
       s_btyp(i) = indexbondtyp(i)
    end do
    s_bk(1:nbondtyp) = para_cbond(1:nbondtyp)
    s_br0(1:nbondtyp) = para_eqbond(1:nbondtyp)
    allocate(s_aijk(max(3*nangl, 1)), s_atyp(max(nangl, 1)))
    allocate(s_ak(max(nangltyp, 1)), s_at0(max(nangltyp, 1)))
    s_aijk = 0; s_atyp = 0; s_ak = 0.0d0; s_at0 = 0.0d0
    do i = 1, nangl
       s_aijk(3*i-2) = iangl(i)
       s_aijk(3*i-1) = jangl(i)
       s_aijk(3*i) = kangl(i)
       s_atyp(i) = indexangltyp(i)
    end do
    s_ak(1:nangltyp) = para_cangl(1:nangltyp)
    s_at0(1:nangltyp) = para_eqangl(1:nangltyp)
    allocate(s_tijkl(max(4*ntors, 1)), s_ttyp(max(ntors, 1)), s_tring(max(ntors, 1)))
    allocate(s_tdiv(max(ntorstyp, 1)), s_t14(max(ntorstyp, 1)), s_tbar(max(ntorstyp, 1)))
    allocate(s_tphs(max(ntorstyp, 1)), s_tper(max(ntorstyp, 1)))
    allocate(s_tdv(max(ntorstyp, 1)), s_tde(max(ntorstyp, 1)))
    s_tijkl = 0; s_ttyp = 0; s_tring = 0.0d0; s_tdiv = 1; s_t14 = 0
    s_tbar = 0.0d0; s_tphs = 0.0d0; s_tper = 0.0d0; s_tdv = 1.0d0; s_tde = 1.0d0
    do i = 1, ntors
       s_tijkl(4*i-3) = itors(i)
       s_tijkl(4*i-2) = jtors(i)
       s_tijkl(4*i-1) = ktors(i)
       s_tijkl(4*i) = ltors(i)
       s_ttyp(i) = indextorstyp(i)
       s_tring(i) = ring14(i)
    end do
    do i = 1, ntorstyp
       s_tdiv(i) = para_divfac(i)
       s_tbar(i) = para_barhig(i)
       s_tphs(i) = para_phsang(i)
       s_tper(i) = para_period(i)
       s_tdv(i) = para_divfac_vdw(i)
       s_tde(i) = para_divfac_elc(i)
       s_t14(i) = merge(1, 0, ifcalnonb14(i))
    end do
    allocate(s_rijkl(max(4*ntorsrb, 1)), s_rtyp(max(ntorsrb, 1)), s_rring(max(ntorsrb, 1)))
    allocate(s_rc(max(6*ntorsrbtyp, 1)), s_r14(max(ntorsrbtyp, 1)))
    allocate(s_rdv(max(ntorsrbtyp, 1)), s_rde(max(ntorsrbtyp, 1)))
    s_rijkl = 0; s_rtyp = 0; s_rring = 0.0d0; s_rc = 0.0d0; s_r14 = 0
    s_rdv = 1.0d0; s_rde = 1.0d0
    do i = 1, ntorsrb
       s_rijkl(4*i-3) = itorsrb(i)
       s_rijkl(4*i-2) = jtorsrb(i)
       s_rijkl(4*i-1) = ktorsrb(i)
       s_rijkl(4*i) = ltorsrb(i)
       s_rtyp(i) = indextorsrbtyp(i)
       s_rring(i) = ring14_rb(i)
    end do
    do i = 1, ntorsrbtyp
       s_rc(6*i-5:6*i) = para_barhigrb(1:6, i)
       s_rdv(i) = para_divfac_vdwrb(i)
       s_rde(i) = para_divfac_elcrb(i)
       s_r14(i) = merge(1, 0, ifcalnonb14_rb(i))
    end do
    s_boxd = (/ xcel, ycel, zcel /)
    d%natom = natom
    d%ntype = nt
    d%charge = c_loc(s_chrg)
    d%atype = c_loc(s_typ)
    d%molid = c_loc(s_mol)
    d%lj_eps = c_loc(s_eps)
    d%lj_rmin = c_loc(s_rmin)
    d%inttyp = c_loc(s_ityp)
    d%nexcl = nlistexcl
    d%excl_index = c_loc(s_eidx)
    d%excl_list = c_loc(s_elst)
    d%nbond = nbond
    d%nbondtyp = nbondtyp
    d%bond_ij = c_loc(s_bij)
    d%bond_type = c_loc(s_btyp)
    d%bond_k = c_loc(s_bk)
    d%bond_r0 = c_loc(s_br0)
    d%nangl = nangl
    d%nangltyp = nangltyp
    d%angl_ijk = c_loc(s_aijk)
    d%angl_type = c_loc(s_atyp)
    d%angl_k = c_loc(s_ak)
    d%angl_theta0 = c_loc(s_at0)
    d%ntors = ntors
    d%ntorstyp = ntorstyp
    d%tors_ijkl = c_loc(s_tijkl)
    d%tors_type = c_loc(s_ttyp)
    d%tors_ring14 = c_loc(s_tring)
    d%tors_divfac = c_loc(s_tdiv)
    d%tors_barhig = c_loc(s_tbar)
    d%tors_phsang = c_loc(s_tphs)
    d%tors_period = c_loc(s_tper)
    d%tors_divfac_vdw = c_loc(s_tdv)
    d%tors_divfac_elc = c_loc(s_tde)
    d%tors_if14 = c_loc(s_t14)
    d%ntorsrb = ntorsrb
    d%ntorsrbtyp = ntorsrbtyp
    d%torsrb_ijkl = c_loc(s_rijkl)
    d%torsrb_type = c_loc(s_rtyp)
    d%torsrb_ring14 = c_loc(s_rring)
    d%torsrb_c = c_loc(s_rc)
    d%torsrb_divfac_vdw = c_loc(s_rdv)
    d%torsrb_divfac_elc = c_loc(s_rde)
    d%torsrb_if14 = c_loc(s_r14)
    d%alpha = alpha
    d%rrcut = rrcut
    d%rcut = rcut
    d%nspltbl = nspltbl
    d%nfft1 = nfft1
    d%nfft2 = nfft2
    d%nfft3 = nfft3
    d%pme_order = pme_order
    d%ifnetqcorrp = 0
    d%box = c_loc(s_boxd)
    call shim_check(pgk_create(d, shim_precision(), shim_ctx), 'pgk_create')
    shim_natom = natom
    shim_box = s_boxd
  end subroutine shim_ensure

  subroutine shim_box_update(xcel, ycel, zcel)
    real(8), intent(in) :: xcel, ycel, zcel
    real(c_double) :: b(3)
    b = (/ xcel, ycel, zcel /)
    if (all(b == shim_box)) return
    call shim_check(pgk_set_box(shim_ctx, b), 'pgk_set_box')
    shim_box = b
  end subroutine shim_box_update

  subroutine shim_positions()
    use md_global
    call shim_check(pgk_set_positions(shim_ctx, atmcor(1:3, 1:natom)), 'pgk_set_positions')
  end subroutine shim_positions

  subroutine shim_free()
    if (allocated(s_chrg)) deallocate(s_chrg, s_typ, s_mol, s_eps, s_rmin, s_ityp, &
         & s_eidx, s_elst, s_bij, s_btyp, s_bk, s_br0, s_aijk, s_atyp, s_ak, s_at0, &
         & s_tijkl, s_ttyp, s_tdiv, s_t14, s_tring, s_tbar, s_tphs, s_tper, s_tdv, s_tde, &
         & s_rijkl, s_rtyp, s_r14, s_rring, s_rc, s_rdv, s_rde)
  end subroutine shim_free

  ! terms of the reference the library does not evaluate must be off
  subroutine shim_refuse_unsupported(ifewald, ifspme, ifcstmnb, ifposres, ifpotbias)
    logical, intent(in) :: ifewald, ifspme, ifcstmnb, ifposres, ifpotbias
    if (ifewald .or. .not. ifspme .or. ifcstmnb .or. ifposres .or. ifpotbias) then
       write(6,*) 'calforce_gpu: the GPU path evaluates SPME electrostatics with LJ and ', &
            &     'bonded terms only (no standard Ewald, custom NB, position restraint ', &
            &     'or bias potential)'
       stop
    end if
  end subroutine shim_refuse_unsupported

  ! the library's energy record / tensors -> calforce's outputs
  subroutine shim_outputs(e, v, pot_tot, pot_nonbon, pot_vdw, pot_elc, pot_ewk, pot_ewnq, &
       &                  pot_vdw14, pot_elc14, pot_bond, pot_angl, pot_tors, pot_torsrb, &
       &                  pot_viri_coul, pot_viri_lj, atm_viri_bond, atm_viri_angl, &
       &                  atm_viri_tors, atm_viri_14, pot_virit_coul, pot_virit_lj, &
       &                  atm_virit_bond, atm_virit_angl, atm_virit_tors, atm_virit_14)
    real(c_double), intent(in) :: e(PGK_NENE), v(9, PGK_NTENS)
    real(8), intent(inout) :: pot_tot, pot_nonbon, pot_vdw, pot_elc, pot_ewk, pot_ewnq
    real(8), intent(inout) :: pot_vdw14, pot_elc14, pot_bond, pot_angl, pot_tors, pot_torsrb
    real(8), intent(inout) :: pot_viri_coul, pot_viri_lj, atm_viri_bond, atm_viri_angl
    real(8), intent(inout) :: atm_viri_tors, atm_viri_14
    real(8), intent(inout) :: pot_virit_coul(:,:), pot_virit_lj(:,:), atm_virit_bond(:,:)
    real(8), intent(inout) :: atm_virit_angl(:,:), atm_virit_tors(:,:), atm_virit_14(:,:)
    pot_tot = e(PGK_E_TOT)
    pot_nonbon = e(PGK_E_NONBON)
    pot_vdw = e(PGK_E_VDW)
    pot_elc = e(PGK_E_ELC)
    pot_ewk = e(PGK_E_EWK)
    pot_ewnq = e(PGK_E_EWNQ)
    pot_vdw14 = e(PGK_E_VDW14)
    pot_elc14 = e(PGK_E_ELC14)
    pot_bond = e(PGK_E_BOND)
    pot_angl = e(PGK_E_ANGL)
    pot_tors = e(PGK_E_TORS)
    pot_torsrb = e(PGK_E_TORSRB)
    pot_viri_coul = e(PGK_V_COUL)
    pot_viri_lj = e(PGK_V_LJ)
    atm_viri_bond = e(PGK_V_BOND)
    atm_viri_angl = e(PGK_V_ANGL)
    atm_viri_tors = e(PGK_V_TORS)
    atm_viri_14 = e(PGK_V_14)
    pot_virit_coul(1:3,1:3) = reshape(v(:,1), (/3, 3/))
    pot_virit_lj(1:3,1:3) = reshape(v(:,2), (/3, 3/))
    atm_virit_bond(1:3,1:3) = reshape(v(:,3), (/3, 3/))
    atm_virit_angl(1:3,1:3) = reshape(v(:,4), (/3, 3/))
    atm_virit_tors(1:3,1:3) = reshape(v(:,5), (/3, 3/))
    atm_virit_14(1:3,1:3) = reshape(v(:,6), (/3, 3/))
  end subroutine shim_outputs
  ! a list requested before the context existed, at the current positions
  subroutine shim_pending_list()
    if (shim_pending_book > 0.0d0) then
       call shim_check(pgk_build_list(shim_ctx, shim_pending_book), 'pgk_build_list')
       shim_pending_book = -1.0d0
       shim_have_list = .true.
    end if
    if (.not. shim_have_list) then
       write(6,*) 'calforce_gpu: calforce called before mklist2a / mklist_cell'
       stop
    end if
  end subroutine shim_pending_list

  ! energies the call does not evaluate keep their values (MTS levels)
  subroutine shim_ene_in(e, pot_tot, pot_nonbon, pot_vdw, pot_elc, pot_ewk, pot_ewnq, &
       &                 pot_vdw14, pot_elc14, pot_bond, pot_angl, pot_tors, pot_torsrb)
    real(c_double), intent(out) :: e(PGK_NENE)
    real(8), intent(in) :: pot_tot, pot_nonbon, pot_vdw, pot_elc, pot_ewk, pot_ewnq
    real(8), intent(in) :: pot_vdw14, pot_elc14, pot_bond, pot_angl, pot_tors, pot_torsrb
    e = 0.0d0
    e(PGK_E_TOT) = pot_tot
    e(PGK_E_NONBON) = pot_nonbon
    e(PGK_E_VDW) = pot_vdw
    e(PGK_E_ELC) = pot_elc
    e(PGK_E_EWK) = pot_ewk
    e(PGK_E_EWNQ) = pot_ewnq
    e(PGK_E_VDW14) = pot_vdw14
    e(PGK_E_ELC14) = pot_elc14
    e(PGK_E_BOND) = pot_bond
    e(PGK_E_ANGL) = pot_angl
    e(PGK_E_TORS) = pot_tors
    e(PGK_E_TORSRB) = pot_torsrb
  end subroutine shim_ene_in
end module pgk_shim_state

! ---- neighbour lists (mklist.F90:8-234, mklist_cell.F90:8-352) ----------
! The reference's drivers build the list at the start of a step and call
! calforce at the same positions.  The context needs calforce's arguments
! (alpha, cutoffs, SPME grid), so a list requested before the first force call
! is built by that call, at the same positions; later requests build at once.
subroutine mklist2a(rcut_book, xcel, ycel, zcel, ifbook)
  implicit none
  real(8), intent(in) :: rcut_book, xcel, ycel, zcel
  logical, intent(in) :: ifbook
  if (.not. ifbook) return
  call shim_list(rcut_book, xcel, ycel, zcel)
end subroutine mklist2a

subroutine mklist_cell(rcut_book, xcel, ycel, zcel)
  implicit none
  real(8), intent(in) :: rcut_book, xcel, ycel, zcel
  call shim_list(rcut_book, xcel, ycel, zcel)
end subroutine mklist_cell

subroutine shim_list(rcut_book, xcel, ycel, zcel)
  use md_global
  use pgk_shim_state
  implicit none
  real(8), intent(in) :: rcut_book, xcel, ycel, zcel
  if (.not. c_associated(shim_ctx) .or. shim_natom /= natom) then
     shim_pending_book = rcut_book
     return
  end if
  call shim_box_update(xcel, ycel, zcel)
  call shim_positions()
  call shim_check(pgk_build_list(shim_ctx, rcut_book), 'pgk_build_list')
  shim_pending_book = -1.0d0
  shim_have_list = .true.
end subroutine shim_list

! ---- calforce (calforce.F90:8-1073; interface_mdtech.F90:358-608) --------
subroutine calforce(mts_flag, force, &
     & for_viri_coul, pot_viri_coul, for_viri_lj, pot_viri_lj, &
     & for_viri_mor, pot_viri_mor, for_viri_sh, pot_viri_sh, &
     & for_viri_rfh, pot_viri_rfh, for_viri_dou, pot_viri_dou, &
     & for_viri_cstmnb, pot_viri_cstmnb, for_viri_cstmnbex, pot_viri_cstmnbex, &
     & atm_viri_bond, atm_viri_angl, atm_viri_tors, atm_viri_14, &
     & pot_virit_coul, pot_virit_lj, pot_virit_mor, pot_virit_sh, pot_virit_rfh, &
     & pot_virit_dou, pot_virit_cstmnb, pot_virit_cstmnbex, &
     & atm_virit_bond, atm_virit_angl, atm_virit_tors, atm_virit_14, &
     & mts_bond, mts_angl, mts_anglub, mts_tors, mts_torsrb, mts_torsim, &
     & mts_vdw, mts_ewr, mts_ewk, mts_vdw14, mts_elc14, &
     & mts_mor, mts_sh, mts_rfh, mts_dou, mts_cnpvw, mts_cstmnb, mts_posres, mts_potbias, &
     & npoly, npolytyp, npoly_mole, npoly_atom, nwater, nmatom, nmatyp, nmatomtyp, &
     & xcel, ycel, zcel, ifewald, alpha, kmax, rrcut, &
     & ifspme, nfft1, nfft2, nfft3, pme_order, eps0, div_factor_14vdw, div_factor_14elc, &
     & rcut, rcutmor, rcutsh, rcutrfhfo, rcutrfhoo, rcutrfhoh, &
     & rcutdouo, rcutindouo, rcutdouh, rcutrpvw, pot_ewc, pot_ewnq, &
     & pot_tot, pot_nonbon, pot_vdw, pot_elc, pot_ewk, pot_vdw14, pot_elc14, &
     & pot_bond, pot_angl, pot_anglub, pot_tors, pot_torsrb, pot_torsim, &
     & pot_mor, pot_sh, pot_rfh, pot_dou, pot_rpvw, pot_vw, pot_cstmnb, pot_cstmnbex, &
     & pot_posres, pot_pbias, ifcalpremole, ifcalpreatom, ifnetqcorrp, netchrgsq, &
     & ifcstmnb, ifposres, ifpotbias, outinterval, pressinterval, recinterval, &
     & nstep_book, md_cont, ncstmnbex, current_step, nstep_short, istep_short)
  use md_global, only: natom
  use pgk_shim_state
  implicit none
  integer, intent(in) :: mts_flag
  integer, intent(in) :: mts_bond, mts_angl, mts_anglub, mts_tors, mts_torsrb, mts_torsim
  integer, intent(in) :: mts_vdw, mts_ewr, mts_ewk, mts_vdw14, mts_elc14
  integer, intent(in) :: mts_mor, mts_sh, mts_rfh, mts_dou, mts_cnpvw, mts_cstmnb
  integer, intent(in) :: mts_posres, mts_potbias
  integer, intent(in) :: npoly, npolytyp, npoly_mole(:), npoly_atom(:)
  integer, intent(in) :: nwater, nmatom, nmatyp, nmatomtyp(:)
  real(8), intent(in) :: xcel, ycel, zcel
  logical, intent(in) :: ifewald
  real(8), intent(in) :: alpha
  integer, intent(in) :: kmax
  real(8), intent(in) :: rrcut
  logical, intent(in) :: ifspme
  integer, intent(in) :: nfft1, nfft2, nfft3, pme_order
  real(8), intent(in) :: eps0, div_factor_14vdw, div_factor_14elc, rcut
  real(8), intent(in) :: rcutmor, rcutsh, rcutrfhfo, rcutrfhoo, rcutrfhoh
  real(8), intent(in) :: rcutdouo, rcutindouo, rcutdouh, rcutrpvw
  logical, intent(in) :: ifcstmnb
  real(8), intent(in) :: pot_ewc
  logical, intent(in) :: ifcalpremole, ifcalpreatom, ifnetqcorrp
  real(8), intent(in) :: netchrgsq
  logical, intent(in) :: ifposres, ifpotbias
  integer, intent(in) :: outinterval, pressinterval, recinterval, nstep_book, current_step
  integer, intent(in) :: nstep_short, istep_short, md_cont, ncstmnbex
  real(8), intent(inout) :: force(:,:)
  real(8), intent(inout) :: pot_tot, pot_nonbon, pot_vdw, pot_elc, pot_ewk, pot_ewnq
  real(8), intent(inout) :: pot_vdw14, pot_elc14, pot_bond, pot_angl, pot_anglub
  real(8), intent(inout) :: pot_tors, pot_torsrb, pot_torsim, pot_mor, pot_sh, pot_rfh
  real(8), intent(inout) :: pot_dou, pot_rpvw, pot_vw, pot_cstmnb, pot_cstmnbex(:)
  real(8), intent(inout) :: pot_posres, pot_pbias
  real(8), intent(inout) :: for_viri_coul(:,:), pot_viri_coul, pot_virit_coul(:,:)
  real(8), intent(inout) :: for_viri_lj(:,:), pot_viri_lj, pot_virit_lj(:,:)
  real(8), intent(inout) :: for_viri_mor(:,:), pot_viri_mor, pot_virit_mor(:,:)
  real(8), intent(inout) :: for_viri_sh(:,:), pot_viri_sh, pot_virit_sh(:,:)
  real(8), intent(inout) :: for_viri_rfh(:,:), pot_viri_rfh, pot_virit_rfh(:,:)
  real(8), intent(inout) :: for_viri_dou(:,:), pot_viri_dou, pot_virit_dou(:,:)
  real(8), intent(inout) :: for_viri_cstmnb(:,:), pot_viri_cstmnb, pot_virit_cstmnb(:,:)
  real(8), intent(inout) :: for_viri_cstmnbex(:,:,:), pot_viri_cstmnbex(:)
  real(8), intent(inout) :: pot_virit_cstmnbex(:,:,:)
  real(8), intent(inout) :: atm_viri_bond, atm_virit_bond(:,:), atm_viri_angl
  real(8), intent(inout) :: atm_virit_angl(:,:), atm_viri_tors, atm_virit_tors(:,:)
  real(8), intent(inout) :: atm_viri_14, atm_virit_14(:,:)

  type(pgk_mts) :: m
  type(pgk_force_out) :: o
  real(c_double), target :: e(PGK_NENE), v(9, PGK_NTENS)
  real(c_double), allocatable, target :: f(:,:), fc(:,:), fl(:,:)

  call shim_refuse_unsupported(ifewald, ifspme, ifcstmnb, ifposres, ifpotbias)
  call shim_ensure(xcel, ycel, zcel, alpha, rrcut, rcut, nfft1, nfft2, nfft3, pme_order)
  call shim_positions()
  call shim_pending_list()
  m%bond = mts_bond
  m%angl = mts_angl
  m%tors = mts_tors
  m%torsrb = mts_torsrb
  m%vdw = mts_vdw
  m%ewk = mts_ewk
  m%nonb14 = mts_vdw14
  allocate(f(3, natom))
  o%force = c_loc(f)
  call shim_ene_in(e, pot_tot, pot_nonbon, pot_vdw, pot_elc, pot_ewk, pot_ewnq, pot_vdw14, &
       &           pot_elc14, pot_bond, pot_angl, pot_tors, pot_torsrb)
  o%ene = c_loc(e)
  o%virt = c_loc(v)
  if (ifcalpremole) then
     allocate(fc(3, natom), fl(3, natom))
     o%for_viri_coul = c_loc(fc)
     o%for_viri_lj = c_loc(fl)
  end if
  call shim_check(pgk_calforce(shim_ctx, int(mts_flag, c_int32_t), m, &
       &                       int(merge(1, 0, ifcalpreatom), c_int32_t), o), 'pgk_calforce')
  force(1:3, 1:natom) = f(1:3, 1:natom)
  if (ifcalpremole) then
     for_viri_coul(1:3, 1:natom) = fc(1:3, 1:natom)
     for_viri_lj(1:3, 1:natom) = fl(1:3, 1:natom)
  end if
  call shim_outputs(e, v, pot_tot, pot_nonbon, pot_vdw, pot_elc, pot_ewk, pot_ewnq, &
       &            pot_vdw14, pot_elc14, pot_bond, pot_angl, pot_tors, pot_torsrb, &
       &            pot_viri_coul, pot_viri_lj, atm_viri_bond, atm_viri_angl, &
       &            atm_viri_tors, atm_viri_14, pot_virit_coul, pot_virit_lj, &
       &            atm_virit_bond, atm_virit_angl, atm_virit_tors, atm_virit_14)
  ! terms of the reference this path does not evaluate are absent: zero
  pot_anglub = 0.0d0; pot_torsim = 0.0d0; pot_mor = 0.0d0; pot_sh = 0.0d0
  pot_rfh = 0.0d0; pot_dou = 0.0d0; pot_rpvw = 0.0d0; pot_vw = 0.0d0
  pot_cstmnb = 0.0d0; pot_posres = 0.0d0; pot_pbias = 0.0d0
  deallocate(f)
  if (allocated(fc)) deallocate(fc, fl)
end subroutine calforce

#if defined(HF)
! ---- calforce_hf (calforce_hf.F90:8-1451; interface_mdtech.F90:609-951) --
subroutine calforce_hf(mts_flag, force, &
     & for_viri_coul, pot_viri_coul, for_viri_lj, pot_viri_lj, &
     & for_viri_mor, pot_viri_mor, for_viri_sh, pot_viri_sh, &
     & for_viri_rfh, pot_viri_rfh, for_viri_dou, pot_viri_dou, &
     & for_viri_cstmnb, pot_viri_cstmnb, for_viri_cstmnbex, pot_viri_cstmnbex, &
     & atm_viri_bond, atm_viri_angl, atm_viri_tors, atm_viri_14, &
     & pot_virit_coul, pot_virit_lj, pot_virit_mor, pot_virit_sh, pot_virit_rfh, &
     & pot_virit_dou, pot_virit_cstmnb, pot_virit_cstmnbex, &
     & atm_virit_bond, atm_virit_angl, atm_virit_tors, atm_virit_14, &
     & pot_bo_atm, pot_an_atm, pot_to_atm, pot_14_atm, pot_elin_atm, pot_ljin_atm, &
     & pot_elc_atm, pot_vdw_atm, pot_mor_atm, pot_sh_atm, pot_rfh_atm, pot_dou_atm, &
     & pot_cstmnb_atm, pot_cstmnbex_atm, &
     & viribot_atm, viriant_atm, viritot_atm, viri14t_atm, virielint_atm, viriljint_atm, &
     & virielct_atm, viriljt_atm, virimort_atm, virisht_atm, virirfht_atm, viridout_atm, &
     & viricstmnbt_atm, viricstmnbext_atm, &
     & mts_bond, mts_angl, mts_anglub, mts_tors, mts_torsrb, mts_torsim, &
     & mts_vdw, mts_ewr, mts_ewk, mts_vdw14, mts_elc14, &
     & mts_mor, mts_sh, mts_rfh, mts_dou, mts_cnpvw, mts_cstmnb, mts_posres, mts_potbias, &
     & npoly, npolytyp, npoly_mole, npoly_atom, nwater, nmatom, nmatyp, nmatomtyp, &
     & xcel, ycel, zcel, ifewald, alpha, kmax, rrcut, &
     & ifspme, nfft1, nfft2, nfft3, pme_order, eps0, div_factor_14vdw, div_factor_14elc, &
     & rcut, rcutmor, rcutsh, rcutrfhfo, rcutrfhoo, rcutrfhoh, &
     & rcutdouo, rcutindouo, rcutdouh, rcutrpvw, pot_ewc, pot_ewnq, &
     & pot_tot, pot_nonbon, pot_vdw, pot_elc, pot_ewk, pot_vdw14, pot_elc14, &
     & pot_bond, pot_angl, pot_anglub, pot_tors, pot_torsrb, pot_torsim, &
     & pot_mor, pot_sh, pot_rfh, pot_dou, pot_rpvw, pot_vw, pot_cstmnb, pot_cstmnbex, &
     & pot_posres, pot_pbias, ifcalpremole, ifcalpreatom, ifnetqcorrp, netchrgsq, &
     & ifcstmnb, ifposres, ifpotbias, ifhfvol, nhfregion, hfzpos1, hfzpos2, &
     & hftyp_atm, molecom, outinterval, pressinterval, heatfinterval, recinterval, &
     & nstep_book, md_cont, ncstmnbex, current_step, nstep_short, istep_short)
  use md_global, only: natom, nmole
  use pgk_shim_state
  implicit none
  integer, intent(in) :: mts_flag
  integer, intent(in) :: mts_bond, mts_angl, mts_anglub, mts_tors, mts_torsrb, mts_torsim
  integer, intent(in) :: mts_vdw, mts_ewr, mts_ewk, mts_vdw14, mts_elc14
  integer, intent(in) :: mts_mor, mts_sh, mts_rfh, mts_dou, mts_cnpvw, mts_cstmnb
  integer, intent(in) :: mts_posres, mts_potbias
  integer, intent(in) :: npoly, npolytyp, npoly_mole(:), npoly_atom(:)
  integer, intent(in) :: nwater, nmatom, nmatyp, nmatomtyp(:)
  real(8), intent(in) :: xcel, ycel, zcel
  logical, intent(in) :: ifewald
  real(8), intent(in) :: alpha
  integer, intent(in) :: kmax
  real(8), intent(in) :: rrcut
  logical, intent(in) :: ifspme
  integer, intent(in) :: nfft1, nfft2, nfft3, pme_order
  real(8), intent(in) :: eps0, div_factor_14vdw, div_factor_14elc, rcut
  real(8), intent(in) :: rcutmor, rcutsh, rcutrfhfo, rcutrfhoo, rcutrfhoh
  real(8), intent(in) :: rcutdouo, rcutindouo, rcutdouh, rcutrpvw
  logical, intent(in) :: ifcstmnb
  real(8), intent(in) :: pot_ewc
  logical, intent(in) :: ifcalpremole, ifcalpreatom, ifnetqcorrp
  real(8), intent(in) :: netchrgsq
  logical, intent(in) :: ifposres, ifpotbias
  logical, intent(in) :: ifhfvol
  integer, intent(in) :: nhfregion
  real(8), intent(in) :: hfzpos1(:), hfzpos2(:)
  integer, intent(in) :: hftyp_atm(:)
  real(8), intent(in) :: molecom(:,:)
  integer, intent(in) :: outinterval, pressinterval, heatfinterval, recinterval
  integer, intent(in) :: nstep_book, current_step
  integer, intent(in) :: nstep_short, istep_short, md_cont, ncstmnbex
  real(8), intent(inout) :: force(:,:)
  real(8), intent(inout) :: pot_tot, pot_nonbon, pot_vdw, pot_elc, pot_ewk, pot_ewnq
  real(8), intent(inout) :: pot_vdw14, pot_elc14, pot_bond, pot_angl, pot_anglub
  real(8), intent(inout) :: pot_tors, pot_torsrb, pot_torsim, pot_mor, pot_sh, pot_rfh
  real(8), intent(inout) :: pot_dou, pot_rpvw, pot_vw, pot_cstmnb, pot_cstmnbex(:)
  real(8), intent(inout) :: pot_posres, pot_pbias
  real(8), intent(inout) :: for_viri_coul(:,:), pot_viri_coul, pot_virit_coul(:,:)
  real(8), intent(inout) :: for_viri_lj(:,:), pot_viri_lj, pot_virit_lj(:,:)
  real(8), intent(inout) :: for_viri_mor(:,:), pot_viri_mor, pot_virit_mor(:,:)
  real(8), intent(inout) :: for_viri_sh(:,:), pot_viri_sh, pot_virit_sh(:,:)
  real(8), intent(inout) :: for_viri_rfh(:,:), pot_viri_rfh, pot_virit_rfh(:,:)
  real(8), intent(inout) :: for_viri_dou(:,:), pot_viri_dou, pot_virit_dou(:,:)
  real(8), intent(inout) :: for_viri_cstmnb(:,:), pot_viri_cstmnb, pot_virit_cstmnb(:,:)
  real(8), intent(inout) :: for_viri_cstmnbex(:,:,:), pot_viri_cstmnbex(:)
  real(8), intent(inout) :: pot_virit_cstmnbex(:,:,:)
  real(8), intent(inout) :: atm_viri_bond, atm_virit_bond(:,:), atm_viri_angl
  real(8), intent(inout) :: atm_virit_angl(:,:), atm_viri_tors, atm_virit_tors(:,:)
  real(8), intent(inout) :: atm_viri_14, atm_virit_14(:,:)
  real(8), intent(inout) :: pot_bo_atm(:), pot_an_atm(:), pot_to_atm(:), pot_14_atm(:)
  real(8), intent(inout) :: pot_elin_atm(:), pot_ljin_atm(:), pot_elc_atm(:), pot_vdw_atm(:)
  real(8), intent(inout) :: pot_mor_atm(:), pot_sh_atm(:), pot_rfh_atm(:), pot_dou_atm(:)
  real(8), intent(inout) :: pot_cstmnb_atm(:), pot_cstmnbex_atm(:,:)
  real(8), intent(inout) :: viribot_atm(:,:,:,:), viriant_atm(:,:,:,:)
  real(8), intent(inout) :: viritot_atm(:,:,:,:), viri14t_atm(:,:,:,:)
  real(8), intent(inout) :: virielint_atm(:,:,:,:), viriljint_atm(:,:,:,:)
  real(8), intent(inout) :: virielct_atm(:,:,:,:), viriljt_atm(:,:,:,:)
  real(8), intent(inout) :: virimort_atm(:,:,:,:), virisht_atm(:,:,:,:)
  real(8), intent(inout) :: virirfht_atm(:,:,:,:), viridout_atm(:,:,:,:)
  real(8), intent(inout) :: viricstmnbt_atm(:,:,:,:), viricstmnbext_atm(:,:,:,:,:)

  type(pgk_mts) :: m
  type(pgk_force_out) :: o
  type(pgk_hf_desc) :: hd
  type(pgk_hf_out) :: ho
  real(c_double), target :: e(PGK_NENE), v(9, PGK_NTENS)
  real(c_double), allocatable, target :: f(:,:), fc(:,:), fl(:,:), pe(:,:), w(:,:,:,:,:)
  real(c_double), allocatable, target :: z1(:), z2(:), com(:,:)
  integer(c_int32_t), allocatable, target :: ht(:)
  integer :: nr

  call shim_refuse_unsupported(ifewald, ifspme, ifcstmnb, ifposres, ifpotbias)
  call shim_ensure(xcel, ycel, zcel, alpha, rrcut, rcut, nfft1, nfft2, nfft3, pme_order)
  call shim_positions()
  call shim_pending_list()
  m%bond = mts_bond
  m%angl = mts_angl
  m%tors = mts_tors
  m%torsrb = mts_torsrb
  m%vdw = mts_vdw
  m%ewk = mts_ewk
  m%nonb14 = mts_vdw14
  nr = nhfregion
  allocate(f(3, natom), pe(natom, PGK_NHF), w(3, 3, natom, nr, PGK_NHF))
  allocate(z1(nr), z2(nr), ht(natom), com(3, max(nmole, 1)))
  z1(1:nr) = hfzpos1(1:nr)
  z2(1:nr) = hfzpos2(1:nr)
  ht(1:natom) = hftyp_atm(1:natom)
  com = 0.0d0
  com(1:3, 1:nmole) = molecom(1:3, 1:nmole)
#if defined(_HF_BULK)
  hd%mode = PGK_HF_BULK
#else
  hd%mode = merge(PGK_HF_VOLUME, PGK_HF_SURFACE, ifhfvol)
#endif
  hd%nregion = nr
  hd%z1 = c_loc(z1)
  hd%z2 = c_loc(z2)
  hd%hftyp = c_loc(ht)
  hd%nmol = nmole
  hd%molecom = c_loc(com)
  o%force = c_loc(f)
  call shim_ene_in(e, pot_tot, pot_nonbon, pot_vdw, pot_elc, pot_ewk, pot_ewnq, pot_vdw14, &
       &           pot_elc14, pot_bond, pot_angl, pot_tors, pot_torsrb)
  o%ene = c_loc(e)
  o%virt = c_loc(v)
  if (ifcalpremole) then
     allocate(fc(3, natom), fl(3, natom))
     o%for_viri_coul = c_loc(fc)
     o%for_viri_lj = c_loc(fl)
  end if
  ho%pot_atm = c_loc(pe)
  ho%viri_atm = c_loc(w)
  call shim_check(pgk_calforce_hf(shim_ctx, int(mts_flag, c_int32_t), m, hd, o, ho), &
       &          'pgk_calforce_hf')
  force(1:3, 1:natom) = f(1:3, 1:natom)
  if (ifcalpremole) then
     for_viri_coul(1:3, 1:natom) = fc(1:3, 1:natom)
     for_viri_lj(1:3, 1:natom) = fl(1:3, 1:natom)
  end if
  call shim_outputs(e, v, pot_tot, pot_nonbon, pot_vdw, pot_elc, pot_ewk, pot_ewnq, &
       &            pot_vdw14, pot_elc14, pot_bond, pot_angl, pot_tors, pot_torsrb, &
       &            pot_viri_coul, pot_viri_lj, atm_viri_bond, atm_viri_angl, &
       &            atm_viri_tors, atm_viri_14, pot_virit_coul, pot_virit_lj, &
       &            atm_virit_bond, atm_virit_angl, atm_virit_tors, atm_virit_14)
  pot_anglub = 0.0d0; pot_torsim = 0.0d0; pot_mor = 0.0d0; pot_sh = 0.0d0
  pot_rfh = 0.0d0; pot_dou = 0.0d0; pot_rpvw = 0.0d0; pot_vw = 0.0d0
  pot_cstmnb = 0.0d0; pot_posres = 0.0d0; pot_pbias = 0.0d0
  ! per-atom classes in the library's order (PGK_HF_BO .. PGK_HF_VDW)
  pot_bo_atm(1:natom) = pe(1:natom, 1)
  pot_an_atm(1:natom) = pe(1:natom, 2)
  pot_to_atm(1:natom) = pe(1:natom, 3)
  pot_14_atm(1:natom) = pe(1:natom, 4)
  pot_elin_atm(1:natom) = pe(1:natom, 5)
  pot_ljin_atm(1:natom) = pe(1:natom, 6)
  pot_elc_atm(1:natom) = pe(1:natom, 7)
  pot_vdw_atm(1:natom) = pe(1:natom, 8)
  pot_mor_atm(1:natom) = 0.0d0
  pot_sh_atm(1:natom) = 0.0d0
  pot_rfh_atm(1:natom) = 0.0d0
  pot_dou_atm(1:natom) = 0.0d0
  pot_cstmnb_atm(1:natom) = 0.0d0
  viribot_atm(:, :, 1:natom, 1:nr) = w(:, :, :, :, 1)
  viriant_atm(:, :, 1:natom, 1:nr) = w(:, :, :, :, 2)
  viritot_atm(:, :, 1:natom, 1:nr) = w(:, :, :, :, 3)
  viri14t_atm(:, :, 1:natom, 1:nr) = w(:, :, :, :, 4)
  virielint_atm(:, :, 1:natom, 1:nr) = w(:, :, :, :, 5)
  viriljint_atm(:, :, 1:natom, 1:nr) = w(:, :, :, :, 6)
  virielct_atm(:, :, 1:natom, 1:nr) = w(:, :, :, :, 7)
  viriljt_atm(:, :, 1:natom, 1:nr) = w(:, :, :, :, 8)
  virimort_atm(:, :, 1:natom, 1:nr) = 0.0d0
  virisht_atm(:, :, 1:natom, 1:nr) = 0.0d0
  virirfht_atm(:, :, 1:natom, 1:nr) = 0.0d0
  viridout_atm(:, :, 1:natom, 1:nr) = 0.0d0
  viricstmnbt_atm(:, :, 1:natom, 1:nr) = 0.0d0
  deallocate(f, pe, w, z1, z2, ht, com)
  if (allocated(fc)) deallocate(fc, fl)
end subroutine calforce_hf
#endif
