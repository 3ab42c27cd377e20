! Fortran energy driver over the GPU solver: the CCENER -> MP2EQN -> CCEQN
! call sequence of relccsd (ccdriv.F:1715-1792, cceqns.F:141-643) with every
! term handed to librelccsd_hip.so through ISO_C_BINDING.
!
! Input  ccin.bin (stream): 8 int32 header (NO, NV, unused, MAXIT, MAXDIM,
!        NTOL, 0, 0), orbital energies, then FOO, FVO, FVV and the classes
!        OOOO, VOOO, VVOO, VOVO, VOVV, VVVV in relccsd layout.  VOVV and VVVV
!        are streamed to the device in column batches, as GETVOVV / GETVVVV
!        read them (ccgetv.F:225-402).
! Output the CCEQN iteration table on unit 6 (format of cceqns.F:6001) and
!        ccout.bin: EMP2, ECCSD, NIT, T1, T2.
program ccdrv_hip
  use relccsd_hip
  implicit none
  integer(c_int) :: hdr(8), rc, nit
  integer :: no, nv, maxit, maxdim, ntol, cls
  integer(c_int64_t) :: noot, nvvt, nvo, rows(0:5), cols(0:5), c0, nb, batch
  real(c_double), allocatable :: eps(:), foo(:), fvo(:), fvv(:), buf(:), t1(:), t2(:)
  integer(c_int), target :: nos(1), nvs(1)
  type(rcc_sym_t) :: sym
  type(c_ptr) :: h
  real(c_double) :: emp2, ecc, rms, diag, conv

  open (30, file='ccin.bin', access='stream', form='unformatted', status='old')
  read (30) hdr
  no = hdr(1); nv = hdr(2); maxit = hdr(4); maxdim = hdr(5); ntol = hdr(6)
  noot = int(no, c_int64_t) * (no - 1) / 2
  nvvt = int(nv, c_int64_t) * (nv - 1) / 2
  nvo = int(no, c_int64_t) * nv
  allocate (eps(no + nv), foo(2 * no * no), fvo(2 * nvo), fvv(2 * nv * nv))
  read (30) eps, foo, fvo, fvv

  nos(1) = no
  nvs(1) = nv
  sym%nrep = 1
  sym%no = c_loc(nos)
  sym%nv = c_loc(nvs)
  sym%frozen = c_null_ptr
  sym%multb = c_null_ptr
  rc = rcc_create(sym, h)
  call rcc_check(h, rc, 'rcc_create')
  call rcc_check(h, rcc_set_options(h, maxdim, maxit, ntol, 0_c_int), 'rcc_set_options')
  call rcc_check(h, rcc_set_fock(h, foo, fvo, fvv, eps), 'rcc_set_fock')

  rows = [noot, nvo, nvvt, nvo, nvo, nvvt]
  cols = [noot, noot, noot, nvo, nvvt, nvvt]
  do cls = 0, 5
    ! column batches of at most 64 MB, the way the sorted files are streamed
    batch = max(1_c_int64_t, (8_c_int64_t * 1024 * 1024) / max(1_c_int64_t, rows(cls)))
    c0 = 0
    do while (c0 < cols(cls))
      nb = min(batch, cols(cls) - c0)
      if (allocated(buf)) deallocate (buf)
      allocate (buf(2 * rows(cls) * nb))
      read (30) buf
      call rcc_check(h, rcc_set_ints(h, int(cls, c_int), 0_c_int, c0, nb, buf), 'rcc_set_ints')
      c0 = c0 + nb
    end do
  end do
  close (30)

  ! MP2EQN (cceqns.F:33-139)
  call rcc_check(h, rcc_mp2(h, emp2), 'MP2EQN')
  write (*, '(a,t40,f25.15)') ' MP2 correlation energy :', emp2

  ! CCEQN (cceqns.F:141-643)
  conv = 10.0d0 ** (-ntol)
  write (*, '(//3x,a,10x,a,20x,a,7x,a)') 'NIT', 'ENERGY', 'RMS', 'T1-DIAGN'
  write (*, 6001) 0, 0.0d0, 1.0d0, 0.0d0
  ecc = 0.0d0
  rms = 1.0d0
  nit = 0
  do while (nit < maxit)
    nit = nit + 1
    call rcc_check(h, rcc_hointm(h), 'HOINTM')
    call rcc_check(h, rcc_hvintm(h), 'HVINTM')
    call rcc_check(h, rcc_hovintm(h), 'HOVINTM')
    call rcc_check(h, rcc_t1eqns(h), 'T1EQNS')
    call rcc_check(h, rcc_gointm(h), 'GOINTM')
    call rcc_check(h, rcc_gvintm(h), 'GVINTM')
    call rcc_check(h, rcc_aintm(h), 'AINTM')
    call rcc_check(h, rcc_hintm(h), 'HINTM')
    call rcc_check(h, rcc_t2eqns(h), 'T2EQNS')
    call rcc_check(h, rcc_bintm(h), 'BINTM')
    call rcc_check(h, rcc_denom(h), 'DENOM')
    call rcc_check(h, rcc_zcore(h), 'ZCORE')
    call rcc_check(h, rcc_diis(h, rms), 'DIISX')
    call rcc_check(h, rcc_energy(h, ecc), 'CCENG')
    call rcc_check(h, rcc_t1diag(h, diag), 'T1DIAG')
    write (*, 6001) nit, ecc, rms, diag
    call rcc_check(h, rcc_accept(h), 'S -> T')
    if (rms < conv) exit
  end do
  if (rms >= conv) write (*, '(//a,f20.15)') ' Requested convergence not reached in CCSD : RMS = ', rms
  write (*, '(a,t40,f25.15)') ' CCSD correlation energy :', ecc

  allocate (t1(2 * nvo), t2(2 * nvvt * noot))
  call rcc_check(h, rcc_get_amplitudes(h, t1, t2), 'GETAMPT')
  open (32, file='ccout.bin', access='stream', form='unformatted', status='replace')
  write (32) emp2, ecc, dble(nit)
  write (32) t1, t2
  close (32)
  call rcc_destroy(h)
6001 format(2x, i3, 5x, 2f20.15, f10.5)
end program ccdrv_hip
