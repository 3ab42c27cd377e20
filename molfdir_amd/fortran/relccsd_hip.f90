! ISO_C_BINDING interface to librelccsd_hip.so (include/relccsd_hip.h).
!
! This is the module a relccsd build USEs to hand its CCSD amplitude path to
! the GPU: CCENER (ccdriv.F:1574) keeps its structure and calls these
! routines where it called MP2EQN / CCEQN (cceqns.F:33, :141).  Counts are
! 64-bit (the reference's INTEGER*4 word counts overflow at 30o/300v);
! complex arrays are passed as the REAL*8 (re,im) pairs relccsd keeps.
module relccsd_hip
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: RCC_OOOO = 0, RCC_VOOO = 1, RCC_VVOO = 2, &
                               RCC_VOVO = 3, RCC_VOVV = 4, RCC_VVVV = 5

  ! layout ids of rcc_layout_block besides the class ids 0..5
  integer(c_int), parameter :: RCC_LAY_T1 = 6, RCC_LAY_T2 = 7, RCC_LAY_FOO = 8, RCC_LAY_FVV = 9

  type, bind(c) :: rcc_sym_t
    integer(c_int) :: nrep
    type(c_ptr) :: no
    type(c_ptr) :: nv
    type(c_ptr) :: frozen
    type(c_ptr) :: multb
  end type rcc_sym_t

  interface
    integer(c_int) function rcc_create(sym, handle) bind(c, name="rcc_create")
      import :: c_int, c_ptr, rcc_sym_t
      type(rcc_sym_t), intent(in) :: sym
      type(c_ptr), intent(out) :: handle
    end function
    subroutine rcc_destroy(h) bind(c, name="rcc_destroy")
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine
    type(c_ptr) function rcc_last_error(h) bind(c, name="rcc_last_error")
      import :: c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function rcc_set_options(h, maxdim, maxit, ntol, prec) &
        bind(c, name="rcc_set_options")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: maxdim, maxit, ntol, prec
    end function
    integer(c_int) function rcc_set_fock(h, foo, fvo, fvv, eps) bind(c, name="rcc_set_fock")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: foo(*), fvo(*), fvv(*), eps(*)
    end function
    integer(c_int) function rcc_set_ints(h, cls, irrep, col0, ncols, buf) &
        bind(c, name="rcc_set_ints")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: cls, irrep
      integer(c_int64_t), value :: col0, ncols
      real(c_double), intent(in) :: buf(*)
    end function
    integer(c_int) function rcc_get_amplitudes(h, t1, t2) bind(c, name="rcc_get_amplitudes")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: t1(*), t2(*)
    end function
    ! start the iterations from these amplitudes (CCEQN's entry, cceqns.F:255-260)
    integer(c_int) function rcc_set_amplitudes(h, t1, t2) bind(c, name="rcc_set_amplitudes")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: t1(*), t2(*)
    end function
    ! one CCEQN iteration (cceqns.F:293-569): terms, DENOM, ZCORE, DIISX, CCENG, S -> T
    integer(c_int) function rcc_iterate(h, nit, ecc, rms) bind(c, name="rcc_iterate")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: nit
      real(c_double), intent(out) :: ecc, rms
    end function
    integer(c_int) function rcc_layout_block(h, lay, block, rows, cols, offset) &
        bind(c, name="rcc_layout_block")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: h
      integer(c_int), value :: lay, block
      integer(c_int64_t), intent(out) :: rows, cols, offset
    end function
    integer(c_int) function rcc_mp2(h, emp2) bind(c, name="rcc_mp2")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: emp2
    end function
    ! CCSD Lambda equations and response density (no live relccsd
    ! counterpart: its Lambda branches are switched off, ccdriv.F:1918-1924)
    integer(c_int) function rcc_lambda(h, nit, rms) bind(c, name="rcc_lambda")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: nit
      real(c_double), intent(out) :: rms
    end function
    integer(c_int) function rcc_get_lambda(h, l1, l2) bind(c, name="rcc_get_lambda")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: l1(*), l2(*)
    end function
    integer(c_int) function rcc_density(h, dens) bind(c, name="rcc_density")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: dens(*)
    end function
    ! integral front end (RDINTS ccints.F:33-569): classes and Fock from the
    ! MOLFDIR files MRCONEE + MDCINT (C strings: end them with c_null_char)
    integer(c_int) function rcc_read_molfdir(h, mrconee, mdcint, nelec, nnelec, e0e1) &
        bind(c, name="rcc_read_molfdir")
      import :: c_int, c_ptr, c_double, c_char
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: mrconee(*), mdcint(*)
      integer(c_int), intent(in) :: nelec(*)
      integer(c_int), value :: nnelec
      real(c_double), intent(out) :: e0e1(3)
    end function
    ! the same with namelist CCSORT's USEOE / NORECMP (GETFM ccints.F:1080-1287)
    integer(c_int) function rcc_read_molfdir_opts(h, mrconee, mdcint, nelec, nnelec, useoe, &
        norecmp, e0e1) bind(c, name="rcc_read_molfdir_opts")
      import :: c_int, c_ptr, c_double, c_char
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: mrconee(*), mdcint(*)
      integer(c_int), intent(in) :: nelec(*)
      integer(c_int), value :: nnelec, useoe, norecmp
      real(c_double), intent(out) :: e0e1(3)
    end function
    integer(c_int) function rcc_get_array(h, which, out, n) bind(c, name="rcc_get_array")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      integer(c_int), value :: which
      real(c_double), intent(out) :: out(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function rcc_get_eps(h, eps) bind(c, name="rcc_get_eps")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: eps(*)
    end function
    ! MCCRES restart file of the handle's amplitudes (PUTAMPT / CCDUMP)
    integer(c_int) function rcc_write_mccres(h, path, imstat, energies) &
        bind(c, name="rcc_write_mccres")
      import :: c_int, c_ptr, c_double, c_char
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), intent(in) :: imstat(30)
      real(c_double), intent(in) :: energies(7)
    end function
    integer(c_int) function rcc_read_mccres(h, path, imstat, energies) &
        bind(c, name="rcc_read_mccres")
      import :: c_int, c_ptr, c_double, c_char
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), intent(out) :: imstat(30)
      real(c_double), intent(out) :: energies(7)
    end function
    ! one-electron properties: RDPROP (ccints.F:569-764) of operator `name`
    ! (C string) into an (no+nv)^2 complex matrix, EXPVAL (ccgrad.F:261-347)
    ! at level 0 (SCF), 1 (unrelaxed CCSD) or 2 (MP2 relaxed), and CCFOPR's
    ! MP2 relaxed density (ccdriv.F:2058-2090)
    integer(c_int) function rcc_read_property(h, mrconee, mdprop, name, nelec, nnelec, prop) &
        bind(c, name="rcc_read_property")
      import :: c_int, c_ptr, c_double, c_char
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: mrconee(*), mdprop(*), name(*)
      integer(c_int), intent(in) :: nelec(*)
      integer(c_int), value :: nnelec
      real(c_double), intent(out) :: prop(*)
    end function
    integer(c_int) function rcc_expectation(h, prop, level, val) bind(c, name="rcc_expectation")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: prop(*)
      integer(c_int), value :: level
      real(c_double), intent(out) :: val(2)
    end function
    integer(c_int) function rcc_mp2_density(h, maxit, maxdim, ntol, nit, rms) &
        bind(c, name="rcc_mp2_density")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: maxit, maxdim, ntol
      integer(c_int), intent(out) :: nit
      real(c_double), intent(out) :: rms
    end function
    ! non-iterative triples of T3CORR (cceqns.F:1336-1863): ET1, ET2, ET3
    integer(c_int) function rcc_t3corr(h, et) bind(c, name="rcc_t3corr")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: et(3)
    end function
    integer(c_int) function rcc_term(h) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function rcc_scalar(h, x) bind(c)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: x
    end function
  end interface

  ! the per-term entry points share one C signature, int f(rcc_t*)
  procedure(rcc_term), bind(c, name="rcc_hointm")  :: rcc_hointm
  procedure(rcc_term), bind(c, name="rcc_hvintm")  :: rcc_hvintm
  procedure(rcc_term), bind(c, name="rcc_hovintm") :: rcc_hovintm
  procedure(rcc_term), bind(c, name="rcc_t1eqns")  :: rcc_t1eqns
  procedure(rcc_term), bind(c, name="rcc_gointm")  :: rcc_gointm
  procedure(rcc_term), bind(c, name="rcc_gvintm")  :: rcc_gvintm
  procedure(rcc_term), bind(c, name="rcc_aintm")   :: rcc_aintm
  procedure(rcc_term), bind(c, name="rcc_hintm")   :: rcc_hintm
  procedure(rcc_term), bind(c, name="rcc_t2eqns")  :: rcc_t2eqns
  procedure(rcc_term), bind(c, name="rcc_bintm")   :: rcc_bintm
  procedure(rcc_term), bind(c, name="rcc_denom")   :: rcc_denom
  procedure(rcc_term), bind(c, name="rcc_zcore")   :: rcc_zcore
  procedure(rcc_term), bind(c, name="rcc_accept")  :: rcc_accept
  procedure(rcc_scalar), bind(c, name="rcc_diis")   :: rcc_diis
  procedure(rcc_scalar), bind(c, name="rcc_energy") :: rcc_energy
  procedure(rcc_scalar), bind(c, name="rcc_t1diag") :: rcc_t1diag

contains

  ! Maps a non-zero return code to an abort, as relccsd's QUIT does (quit.F:33).
  subroutine rcc_check(h, rc, where)
    type(c_ptr), intent(in) :: h
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(rcc_last_error(h), msg, [1024])
    n = 0
    do while (n < 1024)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write (*, '(3a)') ' relccsd_hip error in ', where, ': '
    write (*, '(1024a)') msg(1:n)
    error stop 1
  end subroutine rcc_check

end module relccsd_hip
