!-----------------------------------------------------------------------
! example_calforce.F90 - a Fortran host driving libpgkforce through
! pgk_force_mod: the calls a calforce call site makes (set positions, build
! the list at the bookkeeping cadence, evaluate all terms on MTS_LONG).
!
!   pgk_example_calforce --layout
!       print c_sizeof of every bind(C) type (checked against the C header)
!   pgk_example_calforce <input> <output>
!       input  (stream, native endian): int32 natom, ntype, nexcl, nfft(3),
!              pme_order, nspltbl, precision, ifnetqcorrp; real64 alpha,
!              rrcut, rcut, rcut_book, box(3); charge(natom); int32
!              atype(natom), molid(natom); real64 lj_eps(ntype**2),
!              lj_rmin(ntype**2); int32 inttyp(ntype**2),
!              excl_index(natom+1), excl_list(nexcl); real64 pos(3,natom)
!       output (stream): real64 force(3,natom), ene(PGK_NENE),
!              virt(9*PGK_NTENS)
! Non-bonded terms and exclusions only; the bonded arrays stay null.
!-----------------------------------------------------------------------
program pgk_example_calforce
  use, intrinsic :: iso_c_binding
  use pgk_force_mod
  implicit none

  type(pgk_system_desc) :: desc
  type(pgk_mts) :: mts
  type(pgk_force_out) :: out
  type(pgk_hf_desc) :: hfd
  type(pgk_hf_out) :: hfo
  type(c_ptr) :: ctx
  character(len=4096) :: arg1, arg2
  integer(c_int32_t) :: hdr(10)
  real(c_double) :: par(4)
  real(c_double), target :: box(3)
  real(c_double), allocatable, target :: charge(:), eps(:), rmin(:), pos(:,:), force(:,:)
  real(c_double), target :: ene(PGK_NENE), virt(9 * PGK_NTENS)
  integer(c_int32_t), allocatable, target :: atype(:), molid(:), inttyp(:), xidx(:), xlst(:)
  integer :: natom, ntype, nexcl, u

  call get_command_argument(1, arg1)
  if (trim(arg1) == '--layout') then
     write(6, '(a, i0)') 'pgk_system_desc ', c_sizeof(desc)
     write(6, '(a, i0)') 'pgk_mts ', c_sizeof(mts)
     write(6, '(a, i0)') 'pgk_force_out ', c_sizeof(out)
     write(6, '(a, i0)') 'pgk_hf_desc ', c_sizeof(hfd)
     write(6, '(a, i0)') 'pgk_hf_out ', c_sizeof(hfo)
     write(6, '(a, i0)') 'abi ', pgk_abi_version()
     stop
  end if
  call get_command_argument(2, arg2)

  open(newunit=u, file=trim(arg1), access='stream', form='unformatted', status='old')
  read(u) hdr
  read(u) par, box
  natom = hdr(1)
  ntype = hdr(2)
  nexcl = hdr(3)
  allocate(charge(natom), atype(natom), molid(natom), eps(ntype * ntype), &
       &   rmin(ntype * ntype), inttyp(ntype * ntype), xidx(natom + 1), xlst(nexcl), &
       &   pos(3, natom), force(3, natom))
  read(u) charge, atype, molid, eps, rmin, inttyp, xidx, xlst, pos
  close(u)

  if (pgk_abi_version() /= PGK_ABI) then
     write(6, *) 'libpgkforce ABI version mismatch'
     stop 1
  end if
  desc%natom = natom
  desc%ntype = ntype
  desc%charge = c_loc(charge)
  desc%atype = c_loc(atype)
  desc%molid = c_loc(molid)
  desc%lj_eps = c_loc(eps)
  desc%lj_rmin = c_loc(rmin)
  desc%inttyp = c_loc(inttyp)
  desc%nexcl = nexcl
  desc%excl_index = c_loc(xidx)
  desc%excl_list = c_loc(xlst)
  desc%nfft1 = hdr(4)
  desc%nfft2 = hdr(5)
  desc%nfft3 = hdr(6)
  desc%pme_order = hdr(7)
  desc%nspltbl = hdr(8)
  desc%ifnetqcorrp = hdr(10)
  desc%alpha = par(1)
  desc%rrcut = par(2)
  desc%rcut = par(3)
  desc%box = c_loc(box)

  call pgk_check(pgk_create(desc, hdr(9), ctx), 'pgk_create')
  call pgk_check(pgk_set_box(ctx, box), 'pgk_set_box')
  call pgk_check(pgk_set_positions(ctx, pos), 'pgk_set_positions')
  call pgk_check(pgk_build_list(ctx, par(4)), 'pgk_build_list')
  ene = 0.0d0
  out%force = c_loc(force)
  out%ene = c_loc(ene)
  out%virt = c_loc(virt)
  call pgk_check(pgk_calforce(ctx, PGK_MTS_LONG, mts, 1, out), 'pgk_calforce')
  call pgk_check(pgk_destroy(ctx), 'pgk_destroy')

  open(newunit=u, file=trim(arg2), access='stream', form='unformatted', status='replace')
  write(u) force, ene, virt
  close(u)
end program pgk_example_calforce
