! Fortran energy driver over the GPU solver: the CCENER -> MP2EQN -> CCEQN
! call sequence of relccsd (ccdriv.F:1715-1792, cceqns.F:141-643) with every
! term handed to librelccsd_hip.so through ISO_C_BINDING.
!
! Input  ccin.bin (stream): 8 int32 header (NO, NV, JOB, MAXIT, MAXDIM,
!        NTOL, NREP (0 = C1), FLAGS: 1 LAMBDA, 2 NFROZ (namelist RELCCSD's
!        frozen occupied spinors per irrep, int32, follow the symmetry data;
!        the lowest NFROZ(irp) occupied spinors of each irrep are frozen, as
!        SYMTAB sets IFROZ, ccdriv.F:838-849); JOB 2 adds the triples corrections
!        after CCSD, as CCENER does with DOCCSDT, ccdriv.F:1799-1837; JOB 4
!        sorts the MOLFDIR files MRCONEE + MDCINT of the working directory
!        instead (CCSETI -> RDINTS, C1, NELEC = NO; header word 8 holds
!        namelist CCSORT's options, 1 USEOE + 2 NORECMP; with MRCONEE's BREIT
!        flag MDBINT is sorted and added) and writes E0, E1, ECORE,
!        FOO, FVO, FVV, the six classes and the orbital energies to
!        ccout.bin, then stops; JOB 5 runs CCFOPR's MP2 block instead of
!        CCSD, ccdriv.F:2028-2090: after MP2, the MP2 relaxed density with
!        MAXIT / MAXDIM / NTOL as the Z-vector's MXITGR / MXDIMGR / NTOLGR,
!        then for each operator named at the end of ccin.bin (int32 NEOPER,
!        then NEOPER names of 8 characters) RDPROP from MDPROP (spinors
!        classified by MRCONEE of the working directory, C1) and its SCF and
!        MP2 expectation values, printed as CCFOPR prints them and written
!        to ccout.bin (Re, Im SCF, Re, Im MP2 per operator, then the
!        Z-vector's iterations and RMS)), orbital energies, for NREP > 1 the fermion
!        irrep of every spinor, the occupied count per irrep and MULTB(.,.,1)
!        of the 2 NREP irreps (int32, as in MRCONEE), then FOO, FVO, FVV and
!        the classes OOOO, VOOO, VVOO, VOVO, VOVV, VVVV in relccsd's
!        (symmetry-blocked) layout.  Every class block is streamed to the
!        device in column batches, as GETVOVV / GETVVVV read them
!        (ccgetv.F:225-402).
! Output the CCEQN iteration table on unit 6 (format of cceqns.F:6001) and
!        ccout.bin: EMP2, ECCSD, NIT, T1, T2 (same layout), and the restart
!        file MCCRES (PUTAMPT / CCDUMP layout); with JOB = 2
!        ET1, ET2, ET3 follow (and the T3CORR summary lines of cceqns.F:
!        1799-1812 are printed); with LAMBDA = 1
!        (C1) the Lambda equations are solved after CCSD and NIT, RMS of that
!        solve and the response density (n x n complex, column-major) follow.
program ccdrv_hip
  use relccsd_hip
  implicit none
  integer(c_int) :: hdr(8), rc, nit
  integer :: no, nv, maxit, maxdim, ntol, cls, nrep, i, j, nsp
  integer(c_int64_t) :: rows, cols, off, c0, nb, batch, nfoo, nfvo, nfvv, nt1, nt2
  real(c_double), allocatable :: eps(:), foo(:), fvo(:), fvv(:), buf(:), t1(:), t2(:), dens(:)
  integer(c_int) :: lnit
  real(c_double) :: lrms
  integer(c_int), allocatable, target :: nos(:), nvs(:), multb(:, :), frz(:)
  integer(c_int), allocatable :: nfroz(:)
  integer(c_int), allocatable :: irrsp(:), mult(:, :)
  type(rcc_sym_t) :: sym
  type(c_ptr) :: h
  real(c_double) :: emp2, ecc, rms, diag, conv, et(3), e01(3)
  integer(c_int) :: nel(1), which, imstat(30)
  real(c_double) :: energ(7)
  integer(c_int64_t) :: narr(12:20)
  integer(c_int) :: neoper, iop
  character(len=8), allocatable :: names(:)
  real(c_double), allocatable :: prop(:)
  real(c_double) :: v0(2), v2(2)

  open (30, file='ccin.bin', access='stream', form='unformatted', status='old')
  read (30) hdr
  no = hdr(1); nv = hdr(2); maxit = hdr(4); maxdim = hdr(5); ntol = hdr(6)
  nrep = max(1, hdr(7))
  nsp = no + nv
  allocate (eps(nsp), nos(nrep), nvs(nrep), multb(nrep, nrep))
  read (30) eps
  if (nrep > 1) then
    ! RDSYMI's inputs (ccdriv.F:2978-2988): NO/NV per irrep and the
    ! fermion x fermion -> boson products with the boson offset removed
    allocate (irrsp(nsp), mult(2 * nrep, 2 * nrep))
    read (30) irrsp
    read (30) nos
    read (30) mult
    do i = 1, nrep
      nvs(i) = count(irrsp == i) - nos(i)
    end do
    do j = 1, nrep
      do i = 1, nrep
        multb(i, j) = mult(i, j) - nrep
      end do
    end do
  else
    nos(1) = no
    nvs(1) = nv
  end if
  allocate (nfroz(nrep), frz(max(1, no)))
  nfroz = 0
  if (hdr(3) /= 4 .and. iand(hdr(8), 2_c_int) /= 0) read (30) nfroz
  ! IFROZ: the first NFROZ(irp) occupied spinors of every irrep (irrep-major order)
  frz = 0
  j = 0
  do i = 1, nrep
    if (nfroz(i) < 0 .or. nfroz(i) > nos(i)) stop 'NFROZ exceeds the occupied spinors'
    frz(j + 1:j + nfroz(i)) = 1
    j = j + nos(i)
  end do

  sym%nrep = nrep
  sym%no = c_loc(nos)
  sym%nv = c_loc(nvs)
  sym%frozen = c_null_ptr
  if (any(nfroz > 0)) sym%frozen = c_loc(frz)
  sym%multb = c_null_ptr
  if (nrep > 1) sym%multb = c_loc(multb)
  rc = rcc_create(sym, h)
  call rcc_check(h, rc, 'rcc_create')
  call rcc_check(h, rcc_set_options(h, maxdim, maxit, ntol, 0_c_int), 'rcc_set_options')

  if (hdr(3) == 4) then
    ! integral sort from the MOLFDIR files (CCSETI -> RDINTS -> GETFM)
    close (30)
    nel(1) = no
    ! header word 8: namelist CCSORT's options (1 USEOE + 2 NORECMP)
    call rcc_check(h, rcc_read_molfdir_opts(h, 'MRCONEE'//c_null_char, 'MDCINT'//c_null_char, &
                   nel, 1_c_int, iand(hdr(8), 1_c_int), iand(hdr(8), 2_c_int) / 2_c_int, e01), &
                   'RDINTS')
    write (*, '(a,t40,f25.15)') ' Nuclear repulsion + core energy :', e01(3)
    write (*, '(a,t40,f25.15)') ' Zero order electronic energy :', e01(1)
    write (*, '(a,t40,f25.15)') ' First order electronic energy :', e01(2)
    write (*, '(a,t40,f25.15)') ' SCF energy :', e01(1) + e01(2) + e01(3)
    narr(12) = int(no, c_int64_t) ** 2
    narr(13) = int(nv, c_int64_t) * no
    narr(14) = int(nv, c_int64_t) ** 2
    narr(15) = (int(no, c_int64_t) * (no - 1) / 2) ** 2
    narr(16) = int(nv, c_int64_t) * no * (int(no, c_int64_t) * (no - 1) / 2)
    narr(17) = (int(nv, c_int64_t) * (nv - 1) / 2) * (int(no, c_int64_t) * (no - 1) / 2)
    narr(18) = (int(nv, c_int64_t) * no) ** 2
    narr(19) = int(nv, c_int64_t) * no * (int(nv, c_int64_t) * (nv - 1) / 2)
    narr(20) = (int(nv, c_int64_t) * (nv - 1) / 2) ** 2
    open (32, file='ccout.bin', access='stream', form='unformatted', status='replace')
    write (32) e01
    do which = 12, 20
      if (allocated(buf)) deallocate (buf)
      allocate (buf(2 * max(1_c_int64_t, narr(which))))
      call rcc_check(h, rcc_get_array(h, which, buf, narr(which)), 'rcc_get_array')
      write (32) buf(1:2 * narr(which))
    end do
    call rcc_check(h, rcc_get_eps(h, eps), 'rcc_get_eps')
    write (32) eps
    close (32)
    call rcc_destroy(h)
    stop
  end if

  ! block sizes from the library's SYMTAB restatement (IOO/IVO/IVV totals)
  call rcc_check(h, rcc_layout_block(h, RCC_LAY_FOO, nrep, rows, cols, nfoo), 'layout')
  call rcc_check(h, rcc_layout_block(h, RCC_LAY_T1, nrep, rows, cols, nfvo), 'layout')
  call rcc_check(h, rcc_layout_block(h, RCC_LAY_FVV, nrep, rows, cols, nfvv), 'layout')
  allocate (foo(2 * nfoo), fvo(2 * nfvo), fvv(2 * nfvv))
  read (30) foo, fvo, fvv
  call rcc_check(h, rcc_set_fock(h, foo, fvo, fvv, eps), 'rcc_set_fock')

  do cls = 0, 5
    do i = 0, nrep - 1
      call rcc_check(h, rcc_layout_block(h, int(cls, c_int), int(i, c_int), rows, cols, off), &
                     'layout')
      ! column batches of at most 64 MB, the way the sorted files are streamed
      batch = max(1_c_int64_t, (8_c_int64_t * 1024 * 1024) / max(1_c_int64_t, rows))
      c0 = 0
      do while (c0 < cols)
        nb = min(batch, cols - c0)
        if (allocated(buf)) deallocate (buf)
        allocate (buf(2 * rows * nb))
        read (30) buf
        call rcc_check(h, rcc_set_ints(h, int(cls, c_int), int(i, c_int), c0, nb, buf), &
                       'rcc_set_ints')
        c0 = c0 + nb
      end do
    end do
  end do
  if (hdr(3) == 5) then
    read (30) neoper
    allocate (names(neoper))
    do iop = 1, neoper
      read (30) names(iop)
    end do
  end if
  close (30)

  ! MP2EQN (cceqns.F:33-139)
  call rcc_check(h, rcc_mp2(h, emp2), 'MP2EQN')
  write (*, '(a,t40,f25.15)') ' MP2 correlation energy :', emp2

  if (hdr(3) == 5) then
    ! CCFOPR (ccdriv.F:2028-2090): SCF and MP2 relaxed expectation values
    call rcc_check(h, rcc_mp2_density(h, int(maxit, c_int), int(maxdim, c_int), &
                   int(ntol, c_int), lnit, lrms), 'CCDENSZ')
    write (*, '(a,i4,a,es10.3)') ' Z-vector equations: iterations', lnit, '  RMS', lrms
    allocate (prop(2 * nsp * nsp))
    nel(1) = no
    open (32, file='ccout.bin', access='stream', form='unformatted', status='replace')
    write (*, 1002) '    SCF'
    do iop = 1, neoper
      call rcc_check(h, rcc_read_property(h, 'MRCONEE'//c_null_char, 'MDPROP'//c_null_char, &
                     trim(names(iop))//c_null_char, nel, 1_c_int, prop), 'RDPROP')
      call rcc_check(h, rcc_expectation(h, prop, 0_c_int, v0), 'EXPVAL')
      write (*, 1004) names(iop), v0
    end do
    write (*, 1002) '    MP2'
    do iop = 1, neoper
      call rcc_check(h, rcc_read_property(h, 'MRCONEE'//c_null_char, 'MDPROP'//c_null_char, &
                     trim(names(iop))//c_null_char, nel, 1_c_int, prop), 'RDPROP')
      call rcc_check(h, rcc_expectation(h, prop, 0_c_int, v0), 'EXPVAL')
      call rcc_check(h, rcc_expectation(h, prop, 2_c_int, v2), 'EXPVAL')
      write (*, 1004) names(iop), v2
      write (32) v0, v2
    end do
    write (32) dble(lnit), lrms
    close (32)
    call rcc_destroy(h)
    stop
  end if

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

  call rcc_check(h, rcc_layout_block(h, RCC_LAY_T1, nrep, rows, cols, nt1), 'layout')
  call rcc_check(h, rcc_layout_block(h, RCC_LAY_T2, nrep, rows, cols, nt2), 'layout')
  allocate (t1(2 * nt1), t2(2 * nt2))
  call rcc_check(h, rcc_get_amplitudes(h, t1, t2), 'GETAMPT')
  open (32, file='ccout.bin', access='stream', form='unformatted', status='replace')
  write (32) emp2, ecc, dble(nit)
  write (32) t1, t2
  if (hdr(3) == 2) then
    ! triples corrections on the converged amplitudes (T3CORR)
    call rcc_check(h, rcc_t3corr(h, et), 'T3CORR')
    write (*, '(//a/)') '  Triple excitation estimates'
    write (*, '(a,t40,f25.15)') ' CCSD correlation energy :', ecc
    write (*, '(a,t40,f25.15)') ' 4th order triples correction :', et(1)
    write (*, '(a,t40,f25.15)') ' 5th order triples (T) correction :', et(2)
    write (*, '(a,t40,f25.15)') ' 5th order triples -T  correction :', et(3)
    write (*, '(a,t40,f25.15)') ' Total CCSD+T  energy :', ecc + et(1)
    write (*, '(a,t40,f25.15)') ' Total CCSD(T) energy :', ecc + et(1) + et(2)
    write (*, '(a,t40,f25.15)') ' Total CCSD-T  energy :', ecc + et(1) + et(3)
    write (32) et
  end if
  ! restart file as CCENER leaves it (PUTAMPT + CCDUMP): IMSTAT(4:6) and
  ! ETOT ESCF EMP2 ECCSD ET1 ET2 ET3 (no SCF energy here: ESCF = 0)
  imstat = 0
  imstat(4) = 3
  imstat(5) = 3      ! set after CCEQN whether or not it converged (ccdriv.F:1792)
  energ = 0.0d0
  energ(1) = ecc
  energ(3) = emp2
  energ(4) = ecc
  if (hdr(3) == 2) then
    imstat(6) = 3
    energ(5:7) = et
    energ(1) = ecc + et(1) + et(2)
  end if
  call rcc_check(h, rcc_write_mccres(h, 'MCCRES'//c_null_char, imstat, energ), 'PUTAMPT')
  if (iand(hdr(8), 1_c_int) /= 0) then
    ! Lambda equations and the one-particle response density
    call rcc_check(h, rcc_lambda(h, lnit, lrms), 'LAMBDA')
    write (*, '(a,i4,a,es10.3)') ' Lambda equations: iterations', lnit, '  RMS', lrms
    allocate (dens(2 * nsp * nsp))
    call rcc_check(h, rcc_density(h, dens), 'DENSITY')
    write (32) dble(lnit), lrms
    write (32) dens
  end if
  close (32)
  call rcc_destroy(h)
6001 format(2x, i3, 5x, 2f20.15, f10.5)
1002 format(//, 1x, a7, ' Gradient results', //3x, 'Expectation value', 13x, &
            'Real part', 8x, 'Imaginary part')
1004 format('  <', a8, '> ', t30, 2f20.13)
end program ccdrv_hip
