C  Injection harness for the relccsd CPU reference (test infrastructure).
C
C  relccsd only ingests Kramers-format MDCINT files (ccints.F:765-1029),
C  which the seeded synthetic generator cannot produce.  This driver
C  replaces PROGRAM RELCCSD / CCMAIN (ccmain.F:34, ccdriv.F:375) with a
C  short sequence of calls into the unmodified reference routines:
C
C    RDSYMI (ccdriv.F:2934)  reads the MRCONEE file this program writes
C    CCINPT (ccdriv.F:126)   reads the MOLFDIR.INP file this program writes
C    CCDINI (ccdriv.F:2620), SYMTAB (ccdriv.F:746)
C    RFILE  (ccgenp.F:227)   opens ft40..ft46 as RDINTS would
C    FMTOFILE (ccgetv.F:34) + RWRIT (ccgenp.F:441) store the Fock blocks
C                            and the six integral classes in the layout
C                            GETOOOO..GETVVVV read back (ccgetv.F:89-402)
C
C  Mode 1 then runs CCENER (ccdriv.F:1574): MP2 + CCSD, and writes the
C  final amplitudes.  Mode 2 reads given T1/T2 and calls the per-term
C  routines in the order CCEQN uses them (cceqns.F:313-546), writing
C  every intermediate, so each GPU term can be checked on its own, and
C  ends with CCENG's energy and T1DIAG (cceqns.F:2453) of the new T1.
C  Mode 3 reads given T1/T2 and calls T3CORR (cceqns.F:1336) with the
C  arrays CCENER allocates for it (ccdriv.F:1805-1816), writing the
C  triples corrections ET1, ET2, ET3 (COMMON /ENERG/).
C  Mode 6 reads operator 'PROPTEST' from an MDPROP file written by the
C  caller with RDPROP (ccints.F:569) and evaluates its reference-determinant
C  expectation value with EXPVAL (ccgrad.F:261) on zero densities, as
C  CCGRAD does for its SCF line (ccdriv.F:2028-2039).
C  Mode 7 runs the MP2 relaxed property block of CCFOPR (ccdriv.F:
C  2058-2090) on the given classes: MP2EQN with EQNS = 'LAMBDA',
C  CCDENSC(1), CCDENSZ(1) (the X intermediate and the RPAA Z-vector
C  solve, ccgrad.F:141, ccrpa.F:34) and EXPVAL of operator 'PROPTEST';
C  it writes the MP2 amplitudes, DOO, DVV, -X, the Z-vector (DVO, DOV),
C  the RPA preconditioner DG, the expectation value and the CPU time of
C  the chain (FMTOFILE .. EXPVAL).
C  Mode 5 copies the records of MCCRES.in (a restart file written by the
C  caller) into MCCRES after CCDINI and reads it back with GETAMPT
C  (ccgetv.F:643-709) and the status / energy records CCDUMP writes;
C  mode 1 leaves the MCCRES of its CCENER run in the working directory.
C  Mode 4 takes MRCONEE (records 1-6) and MDCINT as written by the
C  caller into the working directory and runs the integral sort CCSETI
C  (ccdriv.F:1427) -> RDINTS (ccints.F:33) exactly as CCMAIN does with
C  DOSORT (ccdriv.F:555, 638); it writes the Fock blocks (FMTOFILE), the
C  orbital energies after GETFM and the six sorted classes (GETOOOO ..
C  GETVVVV, every irrep).
C
C  Input  : ccin.bin  (stream; header of 8 int32, then float64 data)
C  Output : ccout.bin (stream; float64 data), plus the reference log
C           on unit 6.
C
      PROGRAM CCHRNS
      IMPLICIT INTEGER (A-Z)
      INCLUDE "files.inc"
      INCLUDE "iobuf.inc"
      INCLUDE "diis.inc"
      INCLUDE "inpt.inc"
      INCLUDE "symm.inc"
      INCLUDE "eqns.inc"
      INCLUDE "complex.inc"
      INCLUDE "ccpar.inc"
      INCLUDE "results.inc"
      PARAMETER (MAXSP=1000)
      INTEGER IREPSP(MAXSP*4),IREPSPI(MAXSP*16*2)
      REAL*8 EPS(MAXSP),EPS1(MAXSP)
      INTEGER HDR(8)
      REAL*8, ALLOCATABLE :: CC(:)
      REAL*8, ALLOCATABLE :: FOO(:),FVO(:),FVV(:),V1(:),V2(:),V3(:),
     &                       V4(:),V5(:),V6(:)
      REAL*8, ALLOCATABLE :: T1(:),T2(:),S1(:),S2(:),AZ(:),H(:),
     &                       HO(:),HV(:),HOV(:),GOO(:),GVV(:),
     &                       BUF1(:),BUF2(:),BUF3(:)
      REAL*8 DUM(4),ECC,DIAG,RBUF(RSRCLEN)
      REAL*8, ALLOCATABLE :: PB(:),AVO(:),AOV(:),AOO(:),AVV(:),
     &                       ZD(:)
      COMPLEX*16 EPH,CEXPA
      CHARACTER*8 PNAME
      INTEGER ISZ
      CHARACTER*4 RNAM(32)
      CHARACTER*14 SNAM(16)
      INTEGER MULT(32,32),IRRSP(MAXSP),NELRP(16),NFRZ(16)
      CHARACTER*100 FLINE
      CHARACTER*200 LINE
      LOGICAL BRT,DONE
      REAL TC0,TC1
C
      MASTER = 0
      MYPROC = 0
      NMPROC = 1
      IN = 5
      IW = 6
      INTERFACE = 'MOLFDIR   '
C
C     Header: NO, NV, MODE, MAXIT, MAXDIM, NTOL, NREP (0 = C1), and for
C     mode 4 the CCSORT options (1 USEOE + 2 NORECMP); in the other modes
C     bit 2 of word 8 says that NFROZ (frozen occupied spinors per irrep,
C     namelist RELCCSD) follows the symmetry data.
C     For NREP > 1 the orbital energies are followed by the fermion irrep
C     (1..NREP) of every spinor, the occupied count per irrep and the
C     2NREP x 2NREP table MULTB(.,.,1), all int32.
C
      OPEN (30,FILE='ccin.bin',ACCESS='STREAM',FORM='UNFORMATTED',
     &      STATUS='OLD')
      READ (30) HDR
      NOCC  = HDR(1)
      NVIR  = HDR(2)
      MODE  = HDR(3)
      NREPI = MAX0(1,HDR(7))
      NSP   = NOCC + NVIR
      READ (30) (EPS1(I),I=1,NSP)
      IF (NREPI.GT.1) THEN
         READ (30) (IRRSP(I),I=1,NSP)
         READ (30) (NELRP(I),I=1,NREPI)
         READ (30) ((MULT(I,J),I=1,2*NREPI),J=1,2*NREPI)
      ELSE
         DO I = 1, NSP
            IRRSP(I) = 1
         ENDDO
         NELRP(1) = NOCC
         MULT(1,1) = 2
         MULT(2,1) = 1
         MULT(1,2) = 1
         MULT(2,2) = 2
      ENDIF
      DO I = 1, 16
         NFRZ(I) = 0
      ENDDO
      IF (MODE.NE.4.AND.IAND(HDR(8),2).NE.0)
     &   READ (30) (NFRZ(I),I=1,NREPI)
C
C     MRCONEE records 1-5 exactly as RDSYMI reads them (ccdriv.F:2978-
C     2988); the higher point group is taken equal to the abelian one.
C
      IF (MODE.EQ.4) GOTO 5
      OPEN (MRCONEE,FILE='MRCONEE',FORM='UNFORMATTED',
     &      STATUS='REPLACE')
      BRT = .FALSE.
      DO I = 1, NREPI
         WRITE (SNAM(I),'(A1,I2.2,11X)') 'A',I
         WRITE (RNAM(I),'(1X,A1,I2.2)') 'A',I
         WRITE (RNAM(NREPI+I),'(1X,A1,I2.2)') 'a',I
      ENDDO
      WRITE (MRCONEE) NSP,BRT,0.0D0
      WRITE (MRCONEE) NREPI,(SNAM(I),I=1,NREPI),(NELRP(I),I=1,NREPI)
      WRITE (MRCONEE) NREPI,(RNAM(I),I=1,2*NREPI)
      WRITE (MRCONEE) ((MULT(I,J),I=1,2*NREPI),J=1,2*NREPI)
      WRITE (MRCONEE) (IRRSP(I),IRRSP(I),EPS1(I),I=1,NSP)
      CLOSE (MRCONEE)
C
C     Namelist input read by CCINPT (ccdriv.F:126-373)
C
    5 CONTINUE
      OPEN (31,FILE='MOLFDIR.INP',FORM='FORMATTED',STATUS='REPLACE')
      WRITE (LINE,'(16(I4,A1))') (NELRP(I),',',I=1,NREPI)
      WRITE (FLINE,'(16(I4,A1))') (NFRZ(I),',',I=1,NREPI)
      WRITE (31,'(A,A,A,A,A)') ' &RELCCSD NELEC=',TRIM(LINE),
     &   ' NFROZ=',TRIM(FLINE),
     &   ' CARITH=T, DOSORT=F, DOENER=T, IPRNT=1, TIMING=T /'
      WRITE (31,'(A,I4,A,I4,A,I4,A)') ' &CCENER DOMP2=T, DOCCSD=T,'//
     &   ' DOCCSDT=F, MAXIT=',HDR(4),', MAXDIM=',HDR(5),
     &   ', NTOL=',HDR(6),' /'
C     HDR(8) (mode 4): bit 1 USEOE, bit 2 NORECMP of namelist CCSORT
      IF (MODE.EQ.4.AND.HDR(8).NE.0) WRITE (31,'(A,L1,A,L1,A)')
     &   ' &CCSORT USEOE=',IAND(HDR(8),1).NE.0,', NORECMP=',
     &   IAND(HDR(8),2).NE.0,' /'
      CLOSE (31)
C
C     Set-up sequence of CCMAIN (ccdriv.F:426-529)
C
      CALL XTIME(1,0,' ')
      CALL RDSYMI (NSPR,EPS1,IREPSP)
      CALL CCINPT (IN,IW,MIOBUF)
      IRW = INTOWP(1)
      RCW = 2
      IF (.NOT.CARITH) CALL QUIT('HARNESS NEEDS CARITH=T')
      MINBUF = SEC2I(1)
      MINSEC = IRW * RCW
      CALL CCDINI
      CALL SYMTAB (NELEC,NFROZ,NINACT,NACT,NSP,EPS1,EPS,IREPSP,
     &             IREPSPI,.FALSE.,.TRUE.)
C
C     Read the Fock blocks and the six classes (reference layout)
C
      ALLOCATE (FOO(2*NFOO),FVO(2*NFVO),FVV(2*NFVV))
      IF (MODE.EQ.4) GOTO 7
      ALLOCATE (V1(2*NV1),V2(2*NV2),V3(2*NV3),V4(2*NV4),V5(2*NV5),
     &          V6(2*NV6))
      READ (30) FOO,FVO,FVV,V1,V2,V3,V4,V5,V6
C
      DO K = 0, 6
         CALL RFILE(ITAPT+K,20)
      ENDDO
      CALL FMTOFILE (.TRUE.,FVO,FOO,FVV)
      CALL RWRIT(ITAPT+1,V1,NV1*RCW*IRW,1)
      CALL RWRIT(ITAPT+2,V2,NV2*RCW*IRW,1)
      CALL RWRIT(ITAPT+3,V3,NV3*RCW*IRW,1)
      CALL RWRIT(ITAPT+4,V4,NV4*RCW*IRW,1)
      CALL RWRIT(ITAPT+5,V5,NV5*RCW*IRW,1)
      CALL RWRIT(ITAPT+6,V6,NV6*RCW*IRW,1)
      DEALLOCATE (V1,V2,V3,V4,V5,V6)
C
    7 OPEN (32,FILE='ccout.bin',ACCESS='STREAM',FORM='UNFORMATTED',
     &      STATUS='REPLACE')
C
      IF (MODE.EQ.4) THEN
C
C        Integral sort from MRCONEE + MDCINT (sizes, then the sort)
C
         KFREE = 1
         LFREE = 2000000000
         MXCORR = 0
         CALL CCSETI (1,DUM,KFREE,LFREE,MXCORR,MREM,EPS,NSP,IREPSP,
     &                IREPSPI)
         ALLOCATE (CC(MXCORR+1024))
         KFREE = 1
         LFREE = MXCORR
         MREM = 0
         CALL CCSETI (2,CC,KFREE,LFREE,MXCORR,MREM,EPS,NSP,IREPSP,
     &                IREPSPI)
         CALL FMTOFILE (.FALSE.,FVO,FOO,FVV)
         WRITE (32) FOO,FVO,FVV
         WRITE (32) (EPS(I),I=1,NSP)
         DEALLOCATE (CC)
         ALLOCATE (V1(2*NV1),V2(2*NV2),V3(2*NV3),V4(2*NV4),V5(2*NV5),
     &             V6(2*NV6))
         CALL GETOOOO (V1)
         CALL GETVOOO (V2)
         CALL GETVVOO (V3)
         CALL GETVOVO (V4)
         WRITE (32) V1,V2,V3,V4
         OFF = 1
         DO IRP = 1, NREP
            ISTART = 0
            NINT = 0
            MINT = 2000000000
            CALL GETVOVV (IRP,ISTART,NINT,DONE,V5(OFF),MINT)
            OFF = OFF + 2*NVO(IRP)*NINT
         ENDDO
         OFF = 1
         DO IRP = 1, NREP
            ISTART = 0
            NINT = 0
            MINT = 2000000000
            CALL GETVVVV (IRP,ISTART,NINT,DONE,V6(OFF),MINT)
            OFF = OFF + 2*NVVT(IRP)*NINT
         ENDDO
         WRITE (32) V5,V6
         WRITE (32) ESCF
C
      ELSEIF (MODE.EQ.6) THEN
C
C        RDPROP + EXPVAL (reference determinant) on a caller's MDPROP
C
         ALLOCATE (PB(2*NSP*NSP),AVO(2*NSP*NSP),AOV(2*NSP*NSP),
     &             AOO(2*NFOO),AVV(2*NFVV),ZD(2*NSP*NSP))
         ZD = 0.0D0
         PNAME = 'PROPTEST'
         CALL RDPROP (1,PNAME,NSP,.TRUE.,PB,IREPSPI,EPH,AVO,AOV,
     &                AOO,AVV)
         CALL EXPVAL (ZD,ZD,ZD,ZD,AOO,AVV,AVO,AOV,EPH,CEXPA)
         WRITE (32) (AOO(I),I=1,2*NFOO),(AVV(I),I=1,2*NFVV),
     &              (AVO(I),I=1,2*NFVO),(AOV(I),I=1,2*NFVO)
         WRITE (32) DBLE(CEXPA),DIMAG(CEXPA)
C
      ELSEIF (MODE.EQ.7) THEN
C
C        MP2 relaxed first-order property as CCFOPR's DOMP2G block
C        (ccdriv.F:2028-2090): MP2 lambda = amplitudes, CCDENSC(1),
C        CCDENSZ(1) (X intermediate + RPAA Z-vector), EXPVAL of the
C        caller's MDPROP operator 'PROPTEST'.  HDR(4..6) give the
C        Z-vector's MAXIT, MAXDIM and NTOL (MXITGR, MXDIMGR, NTOLGR).
C
         EQNS = 'LAMBDA'
         MXITGR = HDR(4)
         MXDIMGR = HDR(5)
         NTOLGR = HDR(6)
         MAXIT = MXITGR
         MAXDIM = MXDIMGR
         NEOPER = 1
         NAMEE(1) = 'PROPTEST'
         NBUF1 = MAX0(JOOOO(NREP+1),JVOOO(NREP+1),
     &                JVVOO(NREP+1),JVOVO(NREP+1))
         NBUF2 = NBUF1
         NV5IRP = 1
         NV6IRP = 1
         DO IRP = 1, NREP
            NV5IRP = MAX0(NV5IRP,NVO (IRP)*NVVT(IRP))
            NV6IRP = MAX0(NV6IRP,NVVT(IRP)*NVVT(IRP))
         ENDDO
         NBUF3 = MAX0(NBUF1,NV5IRP,NV6IRP)
         ISZ = MAX0(NBUF1,NBUF3,NDIMT2) + 1024
         ALLOCATE (T1(2*NDIMT1),T2(2*NDIMT2),S1(2*NDIMT1),
     &             S2(2*NDIMT2),BUF1(2*ISZ),BUF2(2*ISZ),BUF3(2*ISZ),
     &             PB(2*NSP*NSP),AVO(2*NSP*NSP),AOV(2*NSP*NSP),
     &             AOO(2*NFOO),AVV(2*NFVV),ZD(24*NDIMX+2*NFOO+2*NFVV
     &             + 2*(MAXDIM+1)*(MAXDIM+2) + MAXDIM + 64))
         ZD = 0.0D0
         PDOO = 1
         PDVV = PDOO + 2*NFOO
         PXVO = PDVV + 2*NFVV
         PXOV = PXVO + 2*NDIMX
         PDVO = PXOV + 2*NDIMX
         PDOV = PDVO + 2*NDIMX
         PD2VO = PDOV + 2*NDIMX
         PD2OV = PD2VO + 2*NDIMX
         PDG = PD2OV + 2*NDIMX
         PCC1 = PDG + 2*NDIMX
         PBB1 = PCC1 + MAXDIM + 8
         PBB2 = PBB1 + (MAXDIM+1)*(MAXDIM+2) + 8
         PNAME = 'PROPTEST'
         CALL CPU_TIME(TC0)
         CALL FMTOFILE (.FALSE.,FVO,FOO,FVV)
         CALL MP2EQN (EPS,FVO,T1,S1,T2,S2)
         CALL CCDENSC (1,T1,S1,T2,S2,BUF1,BUF2,BUF3,ZD(PDOO),ZD(PDVV),
     &                 ZD(PDVO),ZD(PDOV))
         WRITE (32) T1,T2
         WRITE (32) (ZD(I),I=PDOO,PDOO+2*NFOO-1),
     &              (ZD(I),I=PDVV,PDVV+2*NFVV-1)
         CALL CCDENSZ (1,NSP,IREPSPI,EPS,T1,S1,T2,S2,
     &                 BUF1,BUF2,BUF3,NBUF2,NBUF3,PB,
     &                 ZD(PXVO),ZD(PXOV),ZD(PDOO),ZD(PDVV),ZD(PDVO),
     &                 ZD(PDOV),AVO,AOV,ZD(PD2VO),ZD(PD2OV),ZD(PDG),
     &                 ZD(PCC1),ZD(PBB1),ZD(PBB2),10.0D0**(-NTOLGR))
         WRITE (32) (ZD(I),I=PXVO,PXVO+2*NDIMX-1),
     &              (ZD(I),I=PDVO,PDVO+2*NDIMX-1),
     &              (ZD(I),I=PDOV,PDOV+2*NDIMX-1),
     &              (ZD(I),I=PDG,PDG+NDIMX-1)
         CALL RDPROP (1,PNAME,NSP,.TRUE.,PB,IREPSPI,EPH,AVO,AOV,
     &                AOO,AVV)
         CALL EXPVAL (ZD(PDOO),ZD(PDVV),ZD(PDVO),ZD(PDOV),AOO,AVV,AVO,
     &                AOV,EPH,CEXPA)
         CALL CPU_TIME(TC1)
         WRITE (32) DBLE(CEXPA),DIMAG(CEXPA),DBLE(NDIMX),DBLE(TC1-TC0)
C
      ELSEIF (MODE.EQ.5) THEN
C
C        GETAMPT on a restart file written by the caller
C
         ALLOCATE (T1(2*NDIMT1),T2(2*NDIMT2))
         IRL = IRECLEN (RSRCLEN,0,0)
         INQUIRE (FILE='MCCRES.in',SIZE=ISZ)
         NRS = ISZ / (8*RSRCLEN)
         OPEN (33,FILE='MCCRES.in',ACCESS='DIRECT',STATUS='OLD',
     &         RECL=IRL)
         OPEN (MCCRES,FILE='MCCRES',ACCESS='DIRECT',STATUS='OLD',
     &         RECL=IRL)
         DO IR = 1, NRS
            READ (33,REC=IR) RBUF
            WRITE (MCCRES,REC=IR) RBUF
         ENDDO
         CLOSE (33)
         CLOSE (MCCRES)
         CALL GETAMPT (T1,T2)
         OPEN (MCCRES,FILE='MCCRES',ACCESS='DIRECT',STATUS='OLD',
     &         RECL=IRL)
         READ (MCCRES,REC=2) IMSTAT
         READ (MCCRES,REC=3) RBUF
         CLOSE (MCCRES)
         WRITE (32) T1,T2
         WRITE (32) (DBLE(IMSTAT(I)),I=1,30)
         WRITE (32) (RBUF(I),I=1,7)
C
      ELSEIF (MODE.EQ.1) THEN
C
C        MP2 + CCSD through the reference energy driver
C
         KFREE = 1
         LFREE = 2000000000
         MXCORR = 0
         CALL CCENER (1,DUM,KFREE,LFREE,MXCORR,EPS)
         ALLOCATE (CC(MXCORR+1024))
         KFREE = 1
         LFREE = MXCORR
         CALL CCENER (2,CC,KFREE,LFREE,MXCORR,EPS)
         ALLOCATE (T1(2*NDIMT1),T2(2*NDIMT2))
         CALL GETAMPT (T1,T2)
         WRITE (32) EMP2,ECCSD,DBLE(NIT)
         WRITE (32) T1,T2
C
      ELSEIF (MODE.EQ.3) THEN
C
C        Non-iterative triples on given amplitudes
C
         ALLOCATE (T1(2*NDIMT1),T2(2*NDIMT2))
         READ (30) T1,T2
         ALLOCATE (CC(2*NV2 + 2*IVVOOT(NREP+1) + 2*LOVVOT(NREP+1)
     &             + 2*NDIMT2 + 2*KVVVOT(NREP+1) + 6*IVVVTMX
     &             + 2*MAX0(NVVVOTMX,NDIMT2) + 16))
         CC = 0.0D0
         EQNS = 'AMPLTD'
         PVOOO = 1
         PT2S = PVOOO + 2*NV2
         PT2T = PT2S + 2*IVVOOT(NREP+1)
         PVVOO = PT2T + 2*LOVVOT(NREP+1)
         PVVVO = PVVOO + 2*NDIMT2
         PB1 = PVVVO + 2*KVVVOT(NREP+1)
         PB2 = PB1 + 2*IVVVTMX
         PB3 = PB2 + 2*IVVVTMX
         PB4 = PB3 + 2*IVVVTMX
         CALL T3CORR (FVO,EPS,T1,T2,CC(PT2S),CC(PT2T),
     &                CC(PB1),CC(PB2),CC(PB3),CC(PB4),
     &                CC(PVVVO),CC(PVVOO),CC(PVOOO))
         WRITE (32) ET1,ET2,ET3
C
      ELSE
C
C        Per-term evaluation on given amplitudes, CCEQN order
C
         EQNS = 'AMPLTD'
         NBUF1 = MAX0(JOOOO(NREP+1),JVOOO(NREP+1),
     &                JVVOO(NREP+1),JVOVO(NREP+1))
         NBUF2 = NBUF1
         NV5IRP = 1
         NV6IRP = 1
         DO IRP = 1, NREP
            NV5IRP = MAX0(NV5IRP,NVO (IRP)*NVVT(IRP))
            NV6IRP = MAX0(NV6IRP,NVVT(IRP)*NVVT(IRP))
         ENDDO
         NBUF3 = MAX0(NBUF1,NV5IRP,NV6IRP)
         ALLOCATE (T1(2*NDIMT1),T2(2*NDIMT2),S1(2*NDIMT1),
     &             S2(2*NDIMT2),AZ(2*NV1),H(2*NV4),HO(2*NFOO),
     &             HV(2*NFVV),HOV(2*NFVO),GOO(2*NFOO),GVV(2*NFVV),
     &             BUF1(2*NBUF1),BUF2(2*NBUF2),BUF3(2*NBUF3))
         READ (30) T1,T2
         S1 = 0.0D0
         S2 = 0.0D0
         AZ = 0.0D0
         H = 0.0D0
         BUF1 = 0.0D0
         BUF2 = 0.0D0
         BUF3 = 0.0D0
         CALL HOINTM (FOO,T1,T2,BUF1,BUF2,BUF3,HO)
         WRITE (32) HO
         CALL HVINTM (FVV,T1,T2,BUF1,BUF2,BUF3,HV)
         WRITE (32) HV
         CALL HOVINTM (FVO,T1,BUF1,BUF2,HOV)
         WRITE (32) HOV
         CALL T1EQNS (FVO,T1,T2,HO,HV,HOV,H,DUM,GOO,GVV,
     &                BUF1,BUF2,NBUF2,BUF3,NBUF3,S1)
         WRITE (32) S1
         CALL GOINTM (FVO,T1,BUF1,BUF2,HO)
         WRITE (32) HO
         CALL GVINTM (FVO,T1,BUF3,NBUF3,BUF2,NBUF2,HV)
         WRITE (32) HV
         CALL AINTM (T1,T2,BUF1,BUF2,AZ)
         WRITE (32) AZ
         CALL HINTM (T1,T2,BUF1,BUF2,BUF3,NBUF3,H)
         WRITE (32) H
         CALL T2EQNS (T1,T2,HO,HV,HOV,H,AZ,DUM,GOO,GVV,
     &                BUF1,BUF2,BUF3,NBUF3,S2)
         WRITE (32) S2
         CALL BINTM (T1,T2,BUF1,BUF2,NBUF2,BUF3,NBUF3,S2,IW)
         WRITE (32) S2
         CALL DENOM (EPS,S1,S2,S1,S2)
         CALL ZCORE (S1,S2)
         WRITE (32) S1,S2
         CALL CCENG (S1,S2,FVO,BUF1,BUF2,ECC)
         CALL T1DIAG (DIAG,S1,BUF1)
         WRITE (32) ECC,DIAG
      ENDIF
      CLOSE (32)
      CLOSE (30)
      END
