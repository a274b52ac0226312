!-----------------------------------------------------------------------------
! C-callable harness around the reference peachgk_md force path.
!
! Test infrastructure only: it is linked with the reference objects into
! oracle/_ref/libpgkref.so and is used to pin the C restatement in
! oracle/pgk_oracle.c and as the CPU baseline.  It fills the reference's
! module state (md_global, spme_global, mpi_global) from plain arrays, runs the
! reference's own set-up routines (mkexcl, spline_interp_init, fft_pme_init,
! erf_corr_cutoff), list builders (mklist2a, mklist_cell) and force drivers
! (calforce, calforce_hf), and copies the results out.
!
! Calls must run on a thread with a large stack: the reference keeps its PME
! grid and per-atom scratch in automatic arrays (fft_pme.F90:52-60).
!
! Output records (shared with the restatement and the GPU library):
!   ene(0:18): pot_tot pot_nonbon pot_vdw pot_elc pot_ewk pot_ewc pot_ewnq
!              pot_vdw14 pot_elc14 pot_bond pot_angl pot_tors pot_torsrb
!              pot_viri_coul pot_viri_lj atm_viri_bond atm_viri_angl
!              atm_viri_tors atm_viri_14
!   vir(9,6):  pot_virit_coul pot_virit_lj atm_virit_bond atm_virit_angl
!              atm_virit_tors atm_virit_14, each a 3x3 in Fortran order
!-----------------------------------------------------------------------------
module pgkref_harness
  use iso_c_binding
  use md_global
  use spme_global
  use mpi_global
  use interface_mdtech, only: calforce, calforce_hf, mklist2a, mklist_cell, &
       &                      erf_corr_cutoff, calheatf, rattle_c, rattlep_v, nhcint
  use interface_preproc, only: MPI_loopinit, malloc_larray, mkexcl, &
       &                       spline_interp_init, fft_pme_init, prepewk
  use interface_interact, only: ewaldkp_hf
  use interface_timer, only: get_timecount, get_timecount_barrier
  use time_global
  implicit none
  private

  real(8), save :: box_s(3) = 0.0d0
  real(8), save :: alpha_s = 0.0d0, rrcut_s = 0.0d0, rcut_s = 0.0d0
  real(8), save :: pot_ewc_s = 0.0d0, netchrgsq_s = 0.0d0
  integer, save :: nfft_s(3) = 0, order_s = 4
  logical, save :: larray_done = .false.

contains

  subroutine pgkref_init(n, ntyp, box, cor, chrg, typ, mol, welij, radij, &
       &                 inttyp) bind(C, name='pgkref_init')
    integer(c_int), value :: n, ntyp
    real(c_double), intent(in) :: box(3), cor(3,n), chrg(n)
    real(c_double), intent(in) :: welij(ntyp,ntyp), radij(ntyp,ntyp)
    integer(c_int), intent(in) :: typ(n), mol(n), inttyp(ntyp,ntyp)
    integer :: i

    nproc = 1
    irank = 0
    call MPI_loopinit(0)
    if (.not. larray_done) then
       call malloc_larray()
       larray_done = .true.
    end if
    natom = n
    box_s(1:3) = box(1:3)
    atmcor(1:3,1:n) = cor(1:3,1:n)
    atmchrg(1:n) = chrg(1:n)
    atmindex(1:n) = typ(1:n)
    irmolept_list(1:n) = mol(1:n)
    nmole = maxval(mol(1:n))
    vdw_welij(1:ntyp,1:ntyp) = welij(1:ntyp,1:ntyp)
    vdw_radij(1:ntyp,1:ntyp) = radij(1:ntyp,1:ntyp)
    inter_inttyp(1:ntyp,1:ntyp) = inttyp(1:ntyp,1:ntyp)
    fennell_flag = .false.
    nbond = 0
    nangl = 0
    nanglub = 0
    ntors = 0
    ntorsrb = 0
    ntorsim = 0
    netchrgsq_s = 0.0d0
    do i = 1, n
       netchrgsq_s = netchrgsq_s + chrg(i)
    end do
    netchrgsq_s = netchrgsq_s * netchrgsq_s
  end subroutine pgkref_init

  subroutine pgkref_set_box(box) bind(C, name='pgkref_set_box')
    real(c_double), intent(in) :: box(3)
    box_s(1:3) = box(1:3)
  end subroutine pgkref_set_box

  subroutine pgkref_set_positions(cor) bind(C, name='pgkref_set_positions')
    real(c_double), intent(in) :: cor(3,*)
    atmcor(1:3,1:natom) = cor(1:3,1:natom)
  end subroutine pgkref_set_positions

  subroutine pgkref_set_bonds(nb, ib, jb, ityp, ntyp, cb, eqb) &
       &     bind(C, name='pgkref_set_bonds')
    integer(c_int), value :: nb, ntyp
    integer(c_int), intent(in) :: ib(nb), jb(nb), ityp(nb)
    real(c_double), intent(in) :: cb(ntyp), eqb(ntyp)
    nbond = nb
    ibond(1:nb) = ib(1:nb)
    jbond(1:nb) = jb(1:nb)
    indexbondtyp(1:nb) = ityp(1:nb)
    nbondtyp = ntyp
    para_cbond(1:ntyp) = cb(1:ntyp)
    para_eqbond(1:ntyp) = eqb(1:ntyp)
  end subroutine pgkref_set_bonds

  subroutine pgkref_set_angles(na, ia, ja, ka, ityp, ntyp, ca, eqa) &
       &     bind(C, name='pgkref_set_angles')
    integer(c_int), value :: na, ntyp
    integer(c_int), intent(in) :: ia(na), ja(na), ka(na), ityp(na)
    real(c_double), intent(in) :: ca(ntyp), eqa(ntyp)
    nangl = na
    iangl(1:na) = ia(1:na)
    jangl(1:na) = ja(1:na)
    kangl(1:na) = ka(1:na)
    indexangltyp(1:na) = ityp(1:na)
    nangltyp = ntyp
    para_cangl(1:ntyp) = ca(1:ntyp)
    para_eqangl(1:ntyp) = eqa(1:ntyp)
  end subroutine pgkref_set_angles

  subroutine pgkref_set_tors(nt, it, jt, kt, lt, ityp, ring, ntyp, divfac, &
       &                     barhig, phsang, period, dvdw, delc, if14) &
       &     bind(C, name='pgkref_set_tors')
    integer(c_int), value :: nt, ntyp
    integer(c_int), intent(in) :: it(nt), jt(nt), kt(nt), lt(nt), ityp(nt)
    real(c_double), intent(in) :: ring(nt)
    integer(c_int), intent(in) :: divfac(ntyp), if14(ntyp)
    real(c_double), intent(in) :: barhig(ntyp), phsang(ntyp), period(ntyp)
    real(c_double), intent(in) :: dvdw(ntyp), delc(ntyp)
    ntors = nt
    itors(1:nt) = it(1:nt)
    jtors(1:nt) = jt(1:nt)
    ktors(1:nt) = kt(1:nt)
    ltors(1:nt) = lt(1:nt)
    indextorstyp(1:nt) = ityp(1:nt)
    ring14(1:nt) = ring(1:nt)
    ntorstyp = ntyp
    para_divfac(1:ntyp) = divfac(1:ntyp)
    para_barhig(1:ntyp) = barhig(1:ntyp)
    para_phsang(1:ntyp) = phsang(1:ntyp)
    para_period(1:ntyp) = period(1:ntyp)
    para_divfac_vdw(1:ntyp) = dvdw(1:ntyp)
    para_divfac_elc(1:ntyp) = delc(1:ntyp)
    ifcalnonb14(1:ntyp) = (if14(1:ntyp) /= 0)
  end subroutine pgkref_set_tors

  subroutine pgkref_set_torsrb(nt, it, jt, kt, lt, ityp, ring, ntyp, barhig6, &
       &                       dvdw, delc, if14) &
       &     bind(C, name='pgkref_set_torsrb')
    integer(c_int), value :: nt, ntyp
    integer(c_int), intent(in) :: it(nt), jt(nt), kt(nt), lt(nt), ityp(nt)
    real(c_double), intent(in) :: ring(nt)
    real(c_double), intent(in) :: barhig6(6,ntyp), dvdw(ntyp), delc(ntyp)
    integer(c_int), intent(in) :: if14(ntyp)
    ntorsrb = nt
    itorsrb(1:nt) = it(1:nt)
    jtorsrb(1:nt) = jt(1:nt)
    ktorsrb(1:nt) = kt(1:nt)
    ltorsrb(1:nt) = lt(1:nt)
    indextorsrbtyp(1:nt) = ityp(1:nt)
    ring14_rb(1:nt) = ring(1:nt)
    ntorsrbtyp = ntyp
    para_barhigrb(1:6,1:ntyp) = barhig6(1:6,1:ntyp)
    para_divfac_vdwrb(1:ntyp) = dvdw(1:ntyp)
    para_divfac_elcrb(1:ntyp) = delc(1:ntyp)
    ifcalnonb14_rb(1:ntyp) = (if14(1:ntyp) /= 0)
  end subroutine pgkref_set_torsrb

  ! exclusion list from the bonded topology (reference mkexcl.F90)
  function pgkref_mkexcl() bind(C, name='pgkref_mkexcl') result(nl)
    integer(c_int) :: nl
    call mkexcl()
    nl = nlistexcl
  end function pgkref_mkexcl

  subroutine pgkref_get_excl(idx, lst) bind(C, name='pgkref_get_excl')
    integer(c_int), intent(out) :: idx(*), lst(*)
    idx(1:natom+1) = index_excl(1:natom+1)
    lst(1:nlistexcl) = list_excl(1:nlistexcl)
  end subroutine pgkref_get_excl

  subroutine pgkref_set_excl(nl, idx, lst) bind(C, name='pgkref_set_excl')
    integer(c_int), value :: nl
    integer(c_int), intent(in) :: idx(*), lst(nl)
    nlistexcl = nl
    index_excl(1:natom+1) = idx(1:natom+1)
    list_excl(1:nl) = lst(1:nl)
  end subroutine pgkref_set_excl

  subroutine pgkref_init_ewald(alpha, rrcut, rcut, nspl, n1, n2, n3, order, &
       &                       pot_ewc) bind(C, name='pgkref_init_ewald')
    real(c_double), value :: alpha, rrcut, rcut
    integer(c_int), value :: nspl, n1, n2, n3, order
    real(c_double), intent(out) :: pot_ewc
    alpha_s = alpha
    rrcut_s = rrcut
    rcut_s = rcut
    nfft_s(1) = n1
    nfft_s(2) = n2
    nfft_s(3) = n3
    order_s = order
    nspltbl = nspl
    call spline_interp_init(alpha, rrcut)
    call fft_pme_init(n1, n2, n3, order, alpha, pot_ewc_s)
    call erf_corr_cutoff(box_s(1), box_s(2), box_s(3), n1, n2, n3)
    pot_ewc = pot_ewc_s
  end subroutine pgkref_init_ewald

  ! spline tables and SPME moduli as set up by the reference
  subroutine pgkref_get_tables(tbl, bsp, rkc) bind(C, name='pgkref_get_tables')
    real(c_double), intent(out) :: tbl(nspltbl,16), bsp(*), rkc
    tbl(:,1) = spltbl_real(1:nspltbl)
    tbl(:,2) = spl_b_real(1:nspltbl)
    tbl(:,3) = spl_c_real(1:nspltbl)
    tbl(:,4) = spl_d_real(1:nspltbl)
    tbl(:,5) = spltbl_realpot(1:nspltbl)
    tbl(:,6) = spl_b_realpot(1:nspltbl)
    tbl(:,7) = spl_c_realpot(1:nspltbl)
    tbl(:,8) = spl_d_realpot(1:nspltbl)
    tbl(:,9) = spltbl_excs(1:nspltbl)
    tbl(:,10) = spl_b_excs(1:nspltbl)
    tbl(:,11) = spl_c_excs(1:nspltbl)
    tbl(:,12) = spl_d_excs(1:nspltbl)
    tbl(:,13) = spltbl_excspot(1:nspltbl)
    tbl(:,14) = spl_b_excspot(1:nspltbl)
    tbl(:,15) = spl_c_excspot(1:nspltbl)
    tbl(:,16) = spl_d_excspot(1:nspltbl)
    bsp(1:nfft_s(1)) = bsp_mod1(1:nfft_s(1))
    bsp(nfft_s(1)+1:nfft_s(1)+nfft_s(2)) = bsp_mod2(1:nfft_s(2))
    bsp(nfft_s(1)+nfft_s(2)+1:nfft_s(1)+nfft_s(2)+nfft_s(3)) = &
         & bsp_mod3(1:nfft_s(3))
    rkc = rkcut
  end subroutine pgkref_get_tables

  ! neighbour list: mode 0 = mklist2a (O(N^2)), 1 = mklist_cell
  function pgkref_mklist(rcut_book, mode) bind(C, name='pgkref_mklist') &
       &   result(nl)
    real(c_double), value :: rcut_book
    integer(c_int), value :: mode
    integer(c_int) :: nl
    ! the list build is timed as moldyn.F90:1063-1244 times it (table_time)
    call get_timecount_barrier(table_time, table_time_elps, 'START')
    if (mode == 0) then
       call mklist2a(rcut_book, box_s(1), box_s(2), box_s(3), .true.)
    else
       call mklist_cell(rcut_book, box_s(1), box_s(2), box_s(3))
    end if
    call get_timecount(table_time, table_time_elps, 'STOP')
    nl = nlistnball
  end function pgkref_mklist

  ! the reference's own per-term timers (time_global): in a build with
  ! -DTIME_M calforce accumulates drct_time (ewaldr), rcpr_time (SPME) and
  ! bond_time (bonded terms) itself; table_time is the list builds above.
  ! Seconds of process CPU time (etime in the serial build).
  subroutine pgkref_times(t, reset) bind(C, name='pgkref_times')
    real(c_double), intent(out) :: t(5)
    integer(c_int), value :: reset
    t(1) = drct_time
    t(2) = rcpr_time
    t(3) = bond_time
    t(4) = table_time
    t(5) = heatf_time
    if (reset /= 0) then
       drct_time = 0.0d0
       rcpr_time = 0.0d0
       bond_time = 0.0d0
       table_time = 0.0d0
       heatf_time = 0.0d0
    end if
  end subroutine pgkref_times

  subroutine pgkref_get_list(idx, lst) bind(C, name='pgkref_get_list')
    integer(c_int), intent(out) :: idx(*), lst(*)
    idx(1:natom) = nb_indexall(1:natom)
    lst(1:nlistnball) = nb_listall(1:nlistnball)
  end subroutine pgkref_get_list

  subroutine pgkref_set_list(nl, idx, lst) bind(C, name='pgkref_set_list')
    integer(c_int), value :: nl
    integer(c_int), intent(in) :: idx(*), lst(nl)
    nlistnball = nl
    nb_indexall(1:natom) = idx(1:natom)
    nb_listall(1:nl) = lst(1:nl)
  end subroutine pgkref_set_list

  ! One call of calforce (calforce.F90:8-1073) with every active term on the
  ! MTS level given in mts(1:7) = bond angl tors torsrb vdw(=ewald real) ewk 14.
  subroutine pgkref_calforce(mts_flag, mts, ifpreatom, ifnetq, force, ene, &
       &                     vir, fvc, fvl) bind(C, name='pgkref_calforce')
    integer(c_int), value :: mts_flag, ifpreatom, ifnetq
    integer(c_int), intent(in) :: mts(7)
    real(c_double), intent(out) :: force(3,*), ene(0:18), vir(9,6)
    real(c_double), intent(out) :: fvc(3,*), fvl(3,*)

    real(8), allocatable :: f(:,:), fv(:,:,:), fvx(:,:,:)
    real(8) :: pvt(3,3,14), pvx(3,3,1), pv(8), pvex(1)
    real(8) :: p(40), pex(1)
    integer :: npm(1), npa(1), nmt(1)
    logical :: premole, preatom

    allocate(f(3,natom), fv(3,natom,7), fvx(3,natom,1))
    premole = .false.
    preatom = (ifpreatom /= 0)
    npm = 0
    npa = 0
    nmt = 0
    p = 0.0d0
    pex = 0.0d0
    call calforce(mts_flag, f, &
         & fv(:,:,1), pv(1), fv(:,:,2), pv(2), fv(:,:,3), pv(3), &
         & fv(:,:,4), pv(4), fv(:,:,5), pv(5), fv(:,:,6), pv(6), &
         & fv(:,:,7), pv(7), fvx, pvex, &
         & p(20), p(21), p(22), p(23), &
         & pvt(:,:,1), pvt(:,:,2), pvt(:,:,3), pvt(:,:,4), pvt(:,:,5), &
         & pvt(:,:,6), pvt(:,:,7), pvx, &
         & pvt(:,:,8), pvt(:,:,9), pvt(:,:,10), pvt(:,:,11), &
         & mts(1), mts(2), 0, mts(3), mts(4), 0, &
         & mts(5), mts(5), mts(6), mts(7), mts(7), &
         & 0, 0, 0, 0, 0, 0, 0, 0, &
         & 0, 1, npm, npa, 0, 0, 1, nmt, &
         & box_s(1), box_s(2), box_s(3), &
         & .false., alpha_s, 0, rrcut_s, &
         & .true., nfft_s(1), nfft_s(2), nfft_s(3), order_s, &
         & 1.0d0, 2.0d0, 2.0d0, &
         & rcut_s, 0.0d0, 0.0d0, 0.0d0, 0.0d0, 0.0d0, 0.0d0, 0.0d0, 0.0d0, &
         & 0.0d0, &
         & pot_ewc_s, p(7), &
         & p(1), p(2), p(3), p(4), p(5), p(8), p(9), &
         & p(10), p(11), p(24), p(12), p(13), p(25), &
         & p(26), p(27), p(28), p(29), p(30), p(31), &
         & p(32), pex, p(33), p(34), &
         & premole, preatom, (ifnetq /= 0), netchrgsq_s, &
         & .false., .false., .false., &
         & 1, 1, 1, 1, MD_NVE, 0, 1, 1, 1)
    force(1:3,1:natom) = f(1:3,1:natom)
    fvc(1:3,1:natom) = fv(1:3,1:natom,1)
    fvl(1:3,1:natom) = fv(1:3,1:natom,2)
    call pack_ene(p, pv, ene)
    vir(:,1) = reshape(pvt(:,:,1), (/9/))
    vir(:,2) = reshape(pvt(:,:,2), (/9/))
    vir(:,3) = reshape(pvt(:,:,8), (/9/))
    vir(:,4) = reshape(pvt(:,:,9), (/9/))
    vir(:,5) = reshape(pvt(:,:,10), (/9/))
    vir(:,6) = reshape(pvt(:,:,11), (/9/))
    deallocate(f, fv, fvx)
  end subroutine pgkref_calforce

  subroutine pack_ene(p, pv, ene)
    real(8), intent(in) :: p(40), pv(8)
    real(c_double), intent(out) :: ene(0:18)
    ene(0) = p(1)          ! pot_tot
    ene(1) = p(2)          ! pot_nonbon
    ene(2) = p(3)          ! pot_vdw
    ene(3) = p(4)          ! pot_elc
    ene(4) = p(5)          ! pot_ewk
    ene(5) = pot_ewc_s     ! pot_ewc
    ene(6) = p(7)          ! pot_ewnq
    ene(7) = p(8)          ! pot_vdw14
    ene(8) = p(9)          ! pot_elc14
    ene(9) = p(10)         ! pot_bond
    ene(10) = p(11)        ! pot_angl
    ene(11) = p(12)        ! pot_tors
    ene(12) = p(13)        ! pot_torsrb
    ene(13) = pv(1)        ! pot_viri_coul
    ene(14) = pv(2)        ! pot_viri_lj
    ene(15) = p(20)        ! atm_viri_bond
    ene(16) = p(21)        ! atm_viri_angl
    ene(17) = p(22)        ! atm_viri_tors
    ene(18) = p(23)        ! atm_viri_14
  end subroutine pack_ene

  ! One call of calforce_hf (calforce_hf.F90:8-1451) on a heat-flux step.
  ! pe(natom,8): pot_bo pot_an pot_to pot_14 pot_elin pot_ljin pot_elc pot_vdw
  ! wv(3,3,natom,nreg,8): viribot viriant viritot viri14t virielint
  !                       viriljint virielct viriljt
  subroutine pgkref_calforce_hf(mts_flag, mts, ifnetq, ifhfvol, nreg, z1, z2, &
       &                        hftyp, nmol, molecom, force, ene, vir, pe, wv) &
       &     bind(C, name='pgkref_calforce_hf')
    integer(c_int), value :: mts_flag, ifnetq, ifhfvol, nreg, nmol
    integer(c_int), intent(in) :: mts(7), hftyp(*)
    real(c_double), intent(in) :: z1(nreg), z2(nreg), molecom(3,nmol)
    real(c_double), intent(out) :: force(3,*), ene(0:18), vir(9,6)
    real(c_double), intent(out) :: pe(natom,8), wv(3,3,natom,nreg,8)

    real(8), allocatable :: f(:,:), fv(:,:,:), fvx(:,:,:)
    real(8), allocatable :: pa(:,:), pax(:,:), w(:,:,:,:,:), wx(:,:,:,:,:)
    real(8) :: pvt(3,3,14), pvx(3,3,1), pv(8), pvex(1)
    real(8) :: p(40), pex(1)
    integer :: npm(1), npa(1), nmt(1)

    allocate(f(3,natom), fv(3,natom,7), fvx(3,natom,1))
    allocate(pa(natom,13), pax(natom,1))
    allocate(w(3,3,natom,nreg,13), wx(3,3,natom,nreg,1))
    npm = 0
    npa = 0
    nmt = 0
    p = 0.0d0
    pex = 0.0d0
    call calforce_hf(mts_flag, f, &
         & fv(:,:,1), pv(1), fv(:,:,2), pv(2), fv(:,:,3), pv(3), &
         & fv(:,:,4), pv(4), fv(:,:,5), pv(5), fv(:,:,6), pv(6), &
         & fv(:,:,7), pv(7), fvx, pvex, &
         & p(20), p(21), p(22), p(23), &
         & pvt(:,:,1), pvt(:,:,2), pvt(:,:,3), pvt(:,:,4), pvt(:,:,5), &
         & pvt(:,:,6), pvt(:,:,7), pvx, &
         & pvt(:,:,8), pvt(:,:,9), pvt(:,:,10), pvt(:,:,11), &
         & pa(:,1), pa(:,2), pa(:,3), pa(:,4), pa(:,5), pa(:,6), pa(:,7), &
         & pa(:,8), pa(:,9), pa(:,10), pa(:,11), pa(:,12), pa(:,13), pax, &
         & w(:,:,:,:,1), w(:,:,:,:,2), w(:,:,:,:,3), w(:,:,:,:,4), &
         & w(:,:,:,:,5), w(:,:,:,:,6), w(:,:,:,:,7), w(:,:,:,:,8), &
         & w(:,:,:,:,9), w(:,:,:,:,10), w(:,:,:,:,11), w(:,:,:,:,12), &
         & w(:,:,:,:,13), wx, &
         & mts(1), mts(2), 0, mts(3), mts(4), 0, &
         & mts(5), mts(5), mts(6), mts(7), mts(7), &
         & 0, 0, 0, 0, 0, 0, 0, 0, &
         & 0, 1, npm, npa, 0, 0, 1, nmt, &
         & box_s(1), box_s(2), box_s(3), &
         & .false., alpha_s, 0, rrcut_s, &
         & .true., nfft_s(1), nfft_s(2), nfft_s(3), order_s, &
         & 1.0d0, 2.0d0, 2.0d0, &
         & rcut_s, 0.0d0, 0.0d0, 0.0d0, 0.0d0, 0.0d0, 0.0d0, 0.0d0, 0.0d0, &
         & 0.0d0, &
         & pot_ewc_s, p(7), &
         & p(1), p(2), p(3), p(4), p(5), p(8), p(9), &
         & p(10), p(11), p(24), p(12), p(13), p(25), &
         & p(26), p(27), p(28), p(29), p(30), p(31), &
         & p(32), pex, p(33), p(34), &
         & .false., .true., (ifnetq /= 0), netchrgsq_s, &
         & .false., .false., .false., &
         & (ifhfvol /= 0), nreg, z1, z2, hftyp(1:natom), molecom, &
         & 1, 1, 1, 1, 1, MD_HTF, 0, 1, 1, 1)
    force(1:3,1:natom) = f(1:3,1:natom)
    call pack_ene(p, pv, ene)
    vir(:,1) = reshape(pvt(:,:,1), (/9/))
    vir(:,2) = reshape(pvt(:,:,2), (/9/))
    vir(:,3) = reshape(pvt(:,:,8), (/9/))
    vir(:,4) = reshape(pvt(:,:,9), (/9/))
    vir(:,5) = reshape(pvt(:,:,10), (/9/))
    vir(:,6) = reshape(pvt(:,:,11), (/9/))
    pe(:,1:8) = pa(1:natom,1:8)
    wv(:,:,:,:,1:8) = w(:,:,1:natom,1:nreg,1:8)
    deallocate(f, fv, fvx, pa, pax, w, wx)
  end subroutine pgkref_calforce_hf

  ! One call of calheatf (calheatf.F90:8-1470), volume-based (ifhfvol), on
  ! per-atom arrays in pgkref_calforce_hf's layout, all on the long level.
  ! Every molecule is passed as a "poly" molecule (its centre-of-mass
  ! velocity is mass weighted over its atoms, calheatf.F90:548-570).  The
  ! routine stops unless every atom type is 110, 111 or 16
  ! (calheatf.F90:620-630), so the call runs with atmindex set to 16 and
  ! restores it afterwards.
  ! hf(3,18,nreg): kin, pot bo an to 14 elin ljin elc lj, viri (same
  ! order), tot.
  subroutine pgkref_calheatf(nreg, z1, z2, hftyp, nmol, molecom, mass, vel, strm, &
       &                     pe, wv, hf) bind(C, name='pgkref_calheatf')
    integer(c_int), value :: nreg, nmol
    integer(c_int), intent(in) :: hftyp(*)
    real(c_double), intent(in) :: z1(nreg), z2(nreg), molecom(3,nmol)
    real(c_double), intent(in) :: mass(natom), vel(3,natom), strm(3,natom)
    real(c_double), intent(in) :: pe(natom,8), wv(3,3,natom,nreg,8)
    real(c_double), intent(out) :: hf(3,18,nreg)

    integer, parameter :: nex = 3
    real(8), allocatable :: z(:), zx(:,:), w0(:,:,:,:), wx(:,:,:,:,:)
    real(8), allocatable :: hp(:,:,:), hx(:,:,:), hv(:,:,:), hz(:,:)
    real(8), allocatable :: hkx(:,:,:), hpx(:,:,:), hvx(:,:,:), tot(:,:)
    real(8), allocatable :: cor_hf(:,:), com_hf(:,:)
    integer, allocatable :: typ_save(:)
    integer :: i, im, k, c

    allocate(z(natom), zx(natom,nex), w0(3,3,natom,nreg), wx(3,3,natom,nreg,nex))
    allocate(hp(3,nreg,8), hv(3,nreg,8), hz(3,nreg), hx(3,nreg,10))
    allocate(hkx(3,nreg,nex), hpx(3,nreg,nex), hvx(3,nreg,nex), tot(3,nreg))
    allocate(cor_hf(3,natom), com_hf(3,max(nmol,1)), typ_save(natom))
    z = 0.0d0
    zx = 0.0d0
    w0 = 0.0d0
    wx = 0.0d0
    cor_hf = 0.0d0
    com_hf = 0.0d0
    atmmass(1:natom) = mass(1:natom)
    atmvel(1:3,1:natom) = vel(1:3,1:natom)
    atmvel_strm(1:3,1:natom) = strm(1:3,1:natom)
    ! molept lists: the atoms of each molecule in index order
    nmole = maxval(irmolept_list(1:natom))
    k = 1
    do im = 1, nmole
       molept_index(im) = k
       do i = 1, natom
          if (irmolept_list(i) == im) then
             molept_list(k) = i
             k = k + 1
          end if
       end do
    end do
    molept_index(nmole+1) = k
    typ_save(1:natom) = atmindex(1:natom)
    atmindex(1:natom) = 16
    call calheatf(nmole, 0, 0, box_s(1), box_s(2), box_s(3), .true., &
         & nreg, z1, z2, hftyp(1:natom), molecom, &
         & pe(:,1), z, z, pe(:,2), z, z, pe(:,3), z, z, pe(:,4), z, z, &
         & pe(:,5), z, z, pe(:,6), z, z, pe(:,7), z, z, pe(:,8), z, z, &
         & z, z, z, z, z, z, z, z, z, z, z, z, z, z, z, zx, zx, zx, &
         & wv(:,:,:,:,1), w0, w0, wv(:,:,:,:,2), w0, w0, wv(:,:,:,:,3), w0, w0, &
         & wv(:,:,:,:,4), w0, w0, wv(:,:,:,:,5), w0, w0, wv(:,:,:,:,6), w0, w0, &
         & wv(:,:,:,:,7), w0, w0, wv(:,:,:,:,8), w0, w0, &
         & w0, w0, w0, w0, w0, w0, w0, w0, w0, w0, w0, w0, w0, w0, w0, &
         & wx, wx, wx, &
         & hz, hkx, &
         & hp(:,:,1), hp(:,:,2), hp(:,:,3), hp(:,:,4), hp(:,:,5), hp(:,:,6), &
         & hp(:,:,7), hp(:,:,8), hx(:,:,1), hx(:,:,2), hx(:,:,3), hx(:,:,4), &
         & hx(:,:,5), hpx, &
         & hv(:,:,1), hv(:,:,2), hv(:,:,3), hv(:,:,4), hv(:,:,5), hv(:,:,6), &
         & hv(:,:,7), hv(:,:,8), hx(:,:,6), hx(:,:,7), hx(:,:,8), hx(:,:,9), &
         & hx(:,:,10), hvx, &
         & nex, tot, cor_hf, com_hf, 1.0d0, 1, 2)
    atmindex(1:natom) = typ_save(1:natom)
    do k = 1, nreg
       hf(:,1,k) = hz(:,k)
       do c = 1, 8
          hf(:,1+c,k) = hp(:,k,c)
          hf(:,9+c,k) = hv(:,k,c)
       end do
       hf(:,18,k) = tot(:,k)
    end do
    atmvel_strm(1:3,1:natom) = 0.0d0
    deallocate(z, zx, w0, wx, hp, hv, hz, hx, hkx, hpx, hvx, tot, cor_hf, com_hf, typ_save)
  end subroutine pgkref_calheatf

  ! prepewk (prepewk.F90:9-158) then ewaldkp_hf (ewaldk_hf.F90:8-304): the
  ! standard Ewald reciprocal sum with per-atom terms, for the box and alpha
  ! of pgkref_init_ewald.  Adds into force, pvc, virt, pe and wv (the
  ! per-atom virial of region 1); sets pot_ewk.
  subroutine pgkref_ewaldk_hf(kmax, yratio, zratio, ifnetq, force, pot_ewk, pvc, virt, &
       &                      pe, wv) bind(C, name='pgkref_ewaldk_hf')
    integer(c_int), value :: kmax, ifnetq
    real(c_double), value :: yratio, zratio
    real(c_double), intent(inout) :: force(3,natom), pvc, virt(3,3), pe(natom)
    real(c_double), intent(inout) :: wv(3,3,natom)
    real(c_double), intent(out) :: pot_ewk
    real(8), allocatable :: fvc(:,:), w4(:,:,:,:), mc(:,:)
    integer, allocatable :: ht(:)
    real(8) :: ewc, z1(1), z2(1)

    allocate(fvc(3,natom), w4(3,3,natom,1), mc(3,1), ht(natom))
    fvc = 0.0d0
    w4(:,:,:,1) = wv(:,:,:)
    mc = 0.0d0
    ht = 1
    z1 = 0.0d0
    z2 = box_s(3)
    call prepewk(ewc, alpha_s, kmax, box_s(1), box_s(2), box_s(3), yratio, zratio)
    call ewaldkp_hf(box_s(1), box_s(2), box_s(3), alpha_s, force, pot_ewk, &
         &          (ifnetq /= 0), netchrgsq_s, fvc, pvc, virt, pe, w4, &
         &          .true., 1, z1, z2, ht, mc)
    wv(:,:,:) = w4(:,:,:,1)
    deallocate(fvc, w4, mc, ht)
  end subroutine pgkref_ewaldk_hf

  ! ---- integrator pieces (the deterministic parts of the MD step) ----------
  ! Constraints as mkconst leaves them (mkconst.F90:110-164): nconst pairs
  ! (iconst, jconst, dconst), grouped per molecule by index_mconst /
  ! list_mconst with ascending constraint order inside a group; masses and
  ! velocities into atmmass / atmvel (streaming velocities zero).
  subroutine pgkref_set_md(nc, ng, gstart, ic, jc, dc, mass, vel) &
       &     bind(C, name='pgkref_set_md')
    integer(c_int), value :: nc, ng
    integer(c_int), intent(in) :: gstart(ng+1), ic(nc), jc(nc)
    real(c_double), intent(in) :: dc(nc), mass(natom), vel(3,natom)
    integer :: k
    nconst = nc
    iconst(1:nc) = ic(1:nc)
    jconst(1:nc) = jc(1:nc)
    dconst(1:nc) = dc(1:nc)
    index_mconst(:) = 0
    index_mconst(1:ng+1) = gstart(1:ng+1) + 1
    do k = 1, nc
       list_mconst(k) = k
    end do
    atmmass(1:natom) = mass(1:natom)
    atmvel(1:3,1:natom) = vel(1:3,1:natom)
    atmvel_strm(1:3,1:natom) = 0.0d0
  end subroutine pgkref_set_md

  subroutine pgkref_get_md(cor, vel) bind(C, name='pgkref_get_md')
    real(c_double), intent(out) :: cor(3,natom), vel(3,natom)
    cor(1:3,1:natom) = atmcor(1:3,1:natom)
    vel(1:3,1:natom) = atmvel(1:3,1:natom)
  end subroutine pgkref_get_md

  ! rattle_c (rattle.F90:6-272): positions atmcor (already drifted) and the
  ! half-step velocities against the constraints, from the positions at T
  subroutine pgkref_rattle_c(cor_old, eps, dt) bind(C, name='pgkref_rattle_c')
    real(c_double), intent(in) :: cor_old(3,natom)
    real(c_double), value :: eps, dt
    real(8), allocatable :: old(:,:)
    integer :: one
    allocate(old(3,natom))
    old(1:3,1:natom) = cor_old(1:3,1:natom)
    one = 1
    call rattle_c(old, eps, dt, dt, one, one)
    deallocate(old)
  end subroutine pgkref_rattle_c

  ! rattlep_v (rattle_mtk.F90:278-...): velocities against the constraints
  subroutine pgkref_rattlep_v(eps, dt) bind(C, name='pgkref_rattlep_v')
    real(c_double), value :: eps, dt
    real(8) :: vc, vt(3,3)
    vc = 0.0d0
    vt = 0.0d0
    call rattlep_v(eps, dt, dt, 1, 1, 1.0d0, vc, vt)
  end subroutine pgkref_rattlep_v

  ! nhcint (nhcint.F90:8-140) on the chain (qmass, xlogs, vlogs) with the
  ! Yoshida-Suzuki weights w; kT = text (the reference's non-dimensional
  ! temperature), the kinetic energy from atmmass / atmvel
  subroutine pgkref_nhcint(mchain, nyosh, nexti, degfree, dt, text, w, q, x, v) &
       &     bind(C, name='pgkref_nhcint')
    integer(c_int), value :: mchain, nyosh, nexti, degfree
    real(c_double), value :: dt, text
    real(c_double), intent(in) :: w(nyosh), q(mchain)
    real(c_double), intent(inout) :: x(mchain), v(mchain)
    yscoeff(1:nyosh) = w(1:nyosh)
    qmass(1:mchain) = q(1:mchain)
    xlogs(1:mchain) = x(1:mchain)
    vlogs(1:mchain) = v(1:mchain)
    call nhcint(degfree, dt, mchain, text, nexti, nyosh)
    x(1:mchain) = xlogs(1:mchain)
    v(1:mchain) = vlogs(1:mchain)
  end subroutine pgkref_nhcint

end module pgkref_harness
