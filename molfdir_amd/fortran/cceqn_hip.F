C  relccsd's amplitude solvers on the GPU: MP2EQN (cceqns.F:33-139)
C  and CCEQN (cceqns.F:141-643) with relccsd's own names and argument
C  lists, linked into relccsd in place of the two routines of
C  cceqns.F.  relccsd's driver CCENER (ccdriv.F:1715-1792) calls them
C  unchanged; they leave what CCEQN leaves: EMP2, ECCSD and ETOT in
C  COMMON /ENERG/, NIT in /ITER/, the amplitudes in T1/T2 (MP2EQN) or
C  S1/S2 and T1/T2 (CCEQN), and the same printed lines and iteration
C  table (formats of cceqns.F:6001-6079).
C
C  The first call builds one GPU solver handle from SYMTAB's results
C  (COMMON /SYMI/ NREP and MULTB, /SYMD/ NO and NV per irrep, /FREEZE/
C  IFROZ) and streams into device memory, once, the Fock blocks
C  (FMTOFILE, ccgetv.F:34) and the six sorted integral classes
C  (GETOOOO .. GETVOVO whole, GETVOVV and GETVVVV in column batches per
C  irrep, ccgetv.F:89-402).  Every CCEQN iteration then runs on the GPU
C  (terms, DENOM, ZCORE, DIISX, CCENG); only the energies and RMS of
C  the table cross back per iteration.
C
C  Scope: complex arithmetic (CARITH, RCW = 2, the relativistic case
C  this solver is for); CCEQN with EQNS = 'AMPLTD'; MP2EQN with 'AMPLTD'
C  or 'LAMBDA' (CCFOPR's MP2 gradient start).  Anything else stops
C  with QUIT, relccsd's error convention.
      MODULE CCHIPS
      USE, INTRINSIC :: ISO_C_BINDING
      IMPLICIT NONE
C     the solver handle, created by the first MP2EQN or CCEQN call
      TYPE(C_PTR), SAVE :: HCC = C_NULL_PTR
      END MODULE CCHIPS
C
C  Creates the handle and uploads the Fock blocks and the integrals.
      SUBROUTINE CCHLOAD (EPS)
      USE RELCCSD_HIP
      USE CCHIPS
      IMPLICIT INTEGER (A-Z)
      REAL*8 EPS(*)
      INCLUDE "files.inc"
      INCLUDE "symm.inc"
      INCLUDE "freeze.inc"
      INCLUDE "complex.inc"
      TYPE(RCC_SYM_T) SYM
      INTEGER(C_INT), ALLOCATABLE, TARGET :: NOS(:),NVS(:),MTB(:,:),
     &                                       FRZ(:)
      INTEGER(C_INT) RC
      INTEGER(C_INT64_T) ROWS,COLS,OFF,C0,NB,LFOO,LFVO,LFVV,LTOT
      REAL*8, ALLOCATABLE :: XFOO(:),XFVO(:),XFVV(:),BUF(:)
      LOGICAL DONE
C     column batches of at most 8 M complex elements (128 MB)
      PARAMETER (MAXBUF = 8*1024*1024)
C
      IF (C_ASSOCIATED(HCC)) RETURN
      IF (RCW.NE.2) CALL QUIT('GPU CCSD: complex arithmetic only')
      ALLOCATE (NOS(NREP),NVS(NREP),MTB(NREP,NREP))
      NOCC = 0
      DO IRP = 1, NREP
         NOS(IRP) = NO(IRP)
         NVS(IRP) = NV(IRP)
         NOCC = NOCC + NO(IRP)
      ENDDO
C     MULTB(.,.,1) after RDSYMI: fermion x fermion -> boson 1..NREP
      DO JRP = 1, NREP
         DO IRP = 1, NREP
            MTB(IRP,JRP) = MULTB(IRP,JRP,1)
         ENDDO
      ENDDO
C     frozen occupied spinors, irrep by irrep (SYMTAB's IFROZ)
      ALLOCATE (FRZ(MAX(1,NOCC)))
      K = 0
      DO IRP = 1, NREP
         DO I = 1, NO(IRP)
            K = K + 1
            FRZ(K) = 0
            IF (IFROZ(I,IRP).NE.0) FRZ(K) = 1
         ENDDO
      ENDDO
      SYM%NREP = NREP
      SYM%NO = C_LOC(NOS)
      SYM%NV = C_LOC(NVS)
      SYM%FROZEN = C_NULL_PTR
      IF (NFROZEN.GT.0) SYM%FROZEN = C_LOC(FRZ)
      SYM%MULTB = C_NULL_PTR
      IF (NREP.GT.1) SYM%MULTB = C_LOC(MTB)
      RC = RCC_CREATE(SYM,HCC)
      CALL RCC_CHECK(HCC,RC,'rcc_create')
C
C     the Fock blocks as CCENER reads them (FMTOFILE), in the layout
C     SYMTAB and the handle share
      CALL RCC_CHECK(HCC,RCC_LAYOUT_BLOCK(HCC,RCC_LAY_FOO,NREP,ROWS,
     &               COLS,LFOO),'layout')
      CALL RCC_CHECK(HCC,RCC_LAYOUT_BLOCK(HCC,RCC_LAY_T1,NREP,ROWS,
     &               COLS,LFVO),'layout')
      CALL RCC_CHECK(HCC,RCC_LAYOUT_BLOCK(HCC,RCC_LAY_FVV,NREP,ROWS,
     &               COLS,LFVV),'layout')
      IF (LFOO.NE.NFOO.OR.LFVO.NE.NFVO.OR.LFVV.NE.NFVV)
     &   CALL QUIT('GPU CCSD: Fock block sizes differ from SYMTAB')
      ALLOCATE (XFOO(2*MAX(1,NFOO)),XFVO(2*MAX(1,NFVO)),
     &          XFVV(2*MAX(1,NFVV)))
      CALL FMTOFILE (.FALSE.,XFVO,XFOO,XFVV)
      RC = RCC_SET_FOCK(HCC,XFOO,XFVO,XFVV,EPS)
      CALL RCC_CHECK(HCC,RC,'FMTOFILE')
C
C     OOOO, VOOO, VVOO, VOVO: whole classes, one block per irrep
      DO ICLS = 0, 3
         IF (ICLS.EQ.0) LTOT = NV1
         IF (ICLS.EQ.1) LTOT = NV2
         IF (ICLS.EQ.2) LTOT = NV3
         IF (ICLS.EQ.3) LTOT = NV4
         IF (ALLOCATED(BUF)) DEALLOCATE (BUF)
         ALLOCATE (BUF(2*MAX(1_C_INT64_T,LTOT)))
         IF (ICLS.EQ.0) CALL GETOOOO (BUF)
         IF (ICLS.EQ.1) CALL GETVOOO (BUF)
         IF (ICLS.EQ.2) CALL GETVVOO (BUF)
         IF (ICLS.EQ.3) CALL GETVOVO (BUF)
         DO IRP = 1, NREP
            CALL RCC_CHECK(HCC,RCC_LAYOUT_BLOCK(HCC,INT(ICLS,C_INT),
     &                     INT(IRP-1,C_INT),ROWS,COLS,OFF),'layout')
            IF (ROWS*COLS.EQ.0) CYCLE
            IF (2*(OFF+ROWS*COLS).GT.SIZE(BUF,KIND=C_INT64_T))
     &         CALL QUIT('GPU CCSD: class block beyond GETxxxx data')
            RC = RCC_SET_INTS(HCC,INT(ICLS,C_INT),INT(IRP-1,C_INT),
     &                        0_C_INT64_T,COLS,BUF(2*OFF+1))
            CALL RCC_CHECK(HCC,RC,'rcc_set_ints')
         ENDDO
      ENDDO
C
C     VOVV and VVVV: column batches per irrep, as GETVOVV / GETVVVV
C     hand them to CCEQN's terms
      DO ICLS = 4, 5
         DO IRP = 1, NREP
            CALL RCC_CHECK(HCC,RCC_LAYOUT_BLOCK(HCC,INT(ICLS,C_INT),
     &                     INT(IRP-1,C_INT),ROWS,COLS,OFF),'layout')
            IF (ICLS.EQ.4.AND.ROWS.NE.NVO(IRP))
     &         CALL QUIT('GPU CCSD: VOVV batch length differs')
            IF (ICLS.EQ.5.AND.ROWS.NE.NVVT(IRP))
     &         CALL QUIT('GPU CCSD: VVVV batch length differs')
            IF (ROWS*COLS.EQ.0) CYCLE
            MINT = INT(MAX(1_C_INT64_T,MAXBUF/ROWS))
            IF (ALLOCATED(BUF)) DEALLOCATE (BUF)
            ALLOCATE (BUF(2*ROWS*MIN(COLS,INT(MINT,C_INT64_T))))
            ISTART = 0
            DONE = .FALSE.
            DO WHILE (.NOT.DONE)
               IF (ICLS.EQ.4) CALL GETVOVV (IRP,ISTART,NINT,DONE,BUF,
     &                                      MINT)
               IF (ICLS.EQ.5) CALL GETVVVV (IRP,ISTART,NINT,DONE,BUF,
     &                                      MINT)
               IF (NINT.GT.0) THEN
                  C0 = ISTART
                  NB = NINT
                  RC = RCC_SET_INTS(HCC,INT(ICLS,C_INT),
     &                 INT(IRP-1,C_INT),C0,NB,BUF)
                  CALL RCC_CHECK(HCC,RC,'rcc_set_ints')
               ENDIF
               ISTART = ISTART + NINT
               IF (NINT.LE.0.AND..NOT.DONE)
     &            CALL QUIT('GPU CCSD: integral batch read stalled')
            ENDDO
         ENDDO
      ENDDO
      DEALLOCATE (BUF,XFOO,XFVO,XFVV,NOS,NVS,MTB,FRZ)
      RETURN
      END
C
C  MP2EQN (cceqns.F:33-139): T1 = FVO/D, T2 = VVOO/D (frozen spinors
C  zeroed, ZCORE), the MP2 energy and T1 diagnostic; for
C  EQNS = 'LAMBDA' the same T and S = T.
      SUBROUTINE MP2EQN(EPS,FVO,T1,S1,T2,S2)
      USE RELCCSD_HIP
      USE CCHIPS
      IMPLICIT INTEGER (A-Z)
      REAL*8 EPS(*),FVO(*),T1(*),S1(*),T2(*),S2(*)
      INCLUDE "files.inc"
      INCLUDE "param.inc"
      INCLUDE "symm.inc"
      INCLUDE "eqns.inc"
      INCLUDE "inpt.inc"
      INCLUDE "results.inc"
      INCLUDE "complex.inc"
      INCLUDE "ccpar.inc"
      REAL*8 DIAG
      INTEGER(C_INT) RC
C
      IF (EQNS.NE.'AMPLTD'.AND.EQNS.NE.'LAMBDA')
     &   CALL QUIT('GPU MP2EQN: EQNS must be AMPLTD or LAMBDA')
      CALL CCHLOAD (EPS)
      RC = RCC_MP2(HCC,EMP2)
      CALL RCC_CHECK(HCC,RC,'MP2EQN')
      RC = RCC_GET_AMPLITUDES(HCC,T1,T2)
      CALL RCC_CHECK(HCC,RC,'GETAMPT')
      IF (EQNS.EQ.'AMPLTD') THEN
         RC = RCC_T1DIAG(HCC,DIAG)
         CALL RCC_CHECK(HCC,RC,'T1DIAG')
         IF (MYPROC.EQ.MASTER) THEN
            WRITE(IW,6071)
            WRITE(IW,6072) ESCF
            WRITE(IW,6073) EMP2
            ETOT = ESCF + EMP2
            WRITE(IW,6075) ESCF + EMP2
            WRITE(IW,6077) DIAG
         ENDIF
      ELSE
         DO I = 1, RCW*NDIMT1
            S1(I) = T1(I)
         ENDDO
         DO I = 1, RCW*NDIMT2
            S2(I) = T2(I)
         ENDDO
      ENDIF
      RETURN
 6071 FORMAT(//,'  MP2 results',/)
 6072 FORMAT(' SCF energy :',T40,F25.15)
 6073 FORMAT(' MP2 correlation energy :',T40,F25.15)
 6075 FORMAT(' Total MP2 energy :',T40,F25.15)
 6077 FORMAT(' T1 diagnostic :',T40,F25.15)
      END
C
C  CCEQN (cceqns.F:141-643) for the amplitude equations: from the T1/T2
C  CCENER passes, iterations until RMS < CONV or MAXIT (COMMON /ITER/),
C  each one the GPU's HOINTM .. BINTM, DENOM, ZCORE, DIISX (MAXDIM) and
C  CCENG; the work arrays of the CPU solver are not touched.
      SUBROUTINE CCEQN(EPS,FOO,FVO,FVV,T1,S1,T2,S2,AZ,CZ,H,
     &                 HO,HV,HOV,GOO,GVV,
     &                 BUF1,BUF2,BUF3,CC1,BB1,BB2,CONV,
     &                 NBUF2,NBUF3)
      USE RELCCSD_HIP
      USE CCHIPS
      IMPLICIT INTEGER (A-Z)
      REAL*8 EPS(*),FOO(*),FVO(*),FVV(*),T1(*),S1(*),T2(*),S2(*)
      REAL*8 AZ(*),H(*),HO(*),HV(*),HOV(*),BUF1(*),BUF2(*),BUF3(*)
      REAL*8 CZ(*),GOO(*),GVV(*),CC1(*),BB1(*),BB2(*)
      REAL*8 RMS,CONV,DIAG,WALL0,WALL1
      INCLUDE "files.inc"
      INCLUDE "param.inc"
      INCLUDE "symm.inc"
      INCLUDE "diis.inc"
      INCLUDE "eqns.inc"
      INCLUDE "inpt.inc"
      INCLUDE "results.inc"
      INCLUDE "complex.inc"
      INCLUDE "ccpar.inc"
      INTEGER(C_INT) RC,JNIT
      INTEGER(8) CLK0,CLK1,CLKR
C
      IF (EQNS.NE.'AMPLTD')
     &   CALL QUIT('GPU CCEQN: amplitude equations (AMPLTD) only')
      CALL SYSTEM_CLOCK(CLK0,CLKR)
      CALL CCHLOAD (EPS)
      RC = RCC_SET_FOCK(HCC,FOO,FVO,FVV,EPS)
      CALL RCC_CHECK(HCC,RC,'rcc_set_fock')
      NTOLX = NINT(-LOG10(CONV))
      RC = RCC_SET_OPTIONS(HCC,INT(MAXDIM,C_INT),INT(MAXIT,C_INT),
     &                     INT(NTOLX,C_INT),0_C_INT)
      CALL RCC_CHECK(HCC,RC,'rcc_set_options')
      RC = RCC_SET_AMPLITUDES(HCC,T1,T2)
      CALL RCC_CHECK(HCC,RC,'rcc_set_amplitudes')
C
      NIT = 0
      RMS = 1.D0
      ECCSD = 0.D0
      RC = RCC_T1DIAG(HCC,DIAG)
      CALL RCC_CHECK(HCC,RC,'T1DIAG')
      IF (IPRNT.GE.1.AND.MYPROC.EQ.MASTER) THEN
         WRITE(IW,6003)
         WRITE(IW,6001) NIT,ECCSD,RMS,DIAG
      ENDIF
C
 1000 CONTINUE
      NIT = NIT + 1
      IF (NIT.GT.MAXIT) THEN
         NIT = NIT - 1
         IF (MYPROC.EQ.MASTER) WRITE(IW,6004) RMS
         GOTO 2000
      ENDIF
      RC = RCC_ITERATE(HCC,JNIT,ECCSD,RMS)
      CALL RCC_CHECK(HCC,RC,'CCEQN iteration')
      RC = RCC_T1DIAG(HCC,DIAG)
      CALL RCC_CHECK(HCC,RC,'T1DIAG')
      IF (IPRNT.GE.1.AND.MYPROC.EQ.MASTER)
     &   WRITE(IW,6001) NIT,ECCSD,RMS,DIAG
      CALL FLSHFO(IW)
      IF (RMS.LT.CONV) GOTO 2000
      GOTO 1000
C
C     S = T = the last amplitudes, as CCEQN leaves them
 2000 CONTINUE
      RC = RCC_GET_AMPLITUDES(HCC,S1,S2)
      CALL RCC_CHECK(HCC,RC,'GETAMPT')
      DO I = 1, RCW*NDIMT1
         T1(I) = S1(I)
      ENDDO
      DO I = 1, RCW*NDIMT2
         T2(I) = S2(I)
      ENDDO
      IF (MYPROC.EQ.MASTER) THEN
         WRITE(IW,6071)
         ETOT = ESCF + ECCSD
         WRITE(IW,6072) ESCF
         WRITE(IW,6074) ECCSD
         WRITE(IW,6076) ETOT
         WRITE(IW,6077) DIAG
         WRITE(IW,6078) RMS
         WRITE(IW,6079) NIT
      ENDIF
      CALL SYSTEM_CLOCK(CLK1)
      WALL0 = 0.D0
      WALL1 = DBLE(CLK1-CLK0)/DBLE(CLKR)
      IF (TIMING.AND.MYPROC.EQ.MASTER) WRITE(IW,7000) WALL1
      RETURN
 6001 FORMAT(2X,I3,5X,2F20.15,F10.5)
 6003 FORMAT(//3X,'NIT',10X,'ENERGY',20X,'RMS',7X,'T1-DIAGN')
 6004 FORMAT(//' Requested convergence not reached in CCSD : RMS = ',
     &F20.15)
 6071 FORMAT(//,'  CCSD results',/)
 6072 FORMAT(' SCF energy :',T40,F25.15)
 6074 FORMAT(' CCSD correlation energy :',T40,F25.15)
 6076 FORMAT(' Total CCSD energy :',T40,F25.15)
 6077 FORMAT(' T1 diagnostic :',T40,F25.15)
 6078 FORMAT(' Convergence :',T40,F25.15)
 6079 FORMAT(' Number or iterations used :',T60,I5)
 7000 FORMAT(//' Timing of routine CCEQN (GPU solver, wall clock) :'
     &/' Total time in CCEQN :',T30,F12.3,' seconds')
      END
