/* relccsd_hip.h -- C ABI of the MI355X relativistic CCSD amplitude solver.
 *
 * Drop-in for the CCSD amplitude path of MOLFDIR's relccsd: the Fortran
 * driver (CCENER ccdriv.F:1574 -> CCEQN cceqns.F:141) keeps its call
 * sequence and hands its symmetry-sorted complex integral classes and Fock
 * blocks to this library, which holds them in HBM and runs the per-term
 * contractions, denominators, DIIS and energy on the GPU.
 *
 * Conventions (all as relccsd stores them):
 *  - complex numbers are interleaved (re, im) doubles (REAL*8 pairs, RCW=2);
 *  - every array is column-major; class rows are the bra pair, columns the
 *    ket pair; packed pairs are P>Q, enumerated Q outer / P inner;
 *  - VO pairs (A,K) are indexed K*nv + A (A fastest);
 *  - classes: OOOO <KL||IJ>, VOOO <AK||IJ>, VVOO <AB||IJ>, VOVO <AK||CI>,
 *    VOVV <AK||CD>, VVVV <AB||CD>; Fock FOO(K,I), FVO(A,I), FVV(A,C);
 *  - amplitudes T1(A,I) (NDIMT1 = nv*no) and packed T2(AB,IJ) (NDIMT2).
 * Every call returns 0 on success; non-zero means failure and
 * rcc_last_error() describes it (the Fortran shim turns it into QUIT, the
 * reference abort convention quit.F:33).  One handle = one device, one HIP
 * stream; a handle is not shared between host threads.
 */
#ifndef RELCCSD_HIP_H
#define RELCCSD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rcc_handle rcc_t;

/* Integral classes, in the order of the sorted files ft41..ft46
 * (GETOOOO..GETVVVV, ccgetv.F:89-402). */
enum {
    RCC_OOOO = 0, RCC_VOOO = 1, RCC_VVOO = 2, RCC_VOVO = 3, RCC_VOVV = 4, RCC_VVVV = 5
};

/* Intermediates readable with rcc_get_array (for per-term checks). */
enum {
    RCC_ARR_HO = 0, RCC_ARR_HV = 1, RCC_ARR_HOV = 2, RCC_ARR_GO = 3, RCC_ARR_GV = 4,
    RCC_ARR_AZ = 5, RCC_ARR_H = 6, RCC_ARR_S1 = 7, RCC_ARR_S2 = 8, RCC_ARR_T1 = 9,
    RCC_ARR_T2 = 10, RCC_ARR_TAU = 11,
    /* the Fock blocks and the six classes as the handle holds them (for
     * checking the integral front end; symmetry-blocked layout for nrep > 1) */
    RCC_ARR_FOO = 12, RCC_ARR_FVO = 13, RCC_ARR_FVV = 14, RCC_ARR_OOOO = 15,
    RCC_ARR_VOOO = 16, RCC_ARR_VVOO = 17, RCC_ARR_VOVO = 18, RCC_ARR_VOVV = 19,
    RCC_ARR_VVVV = 20,
    /* after rcc_mp2_density: the density blocks DOO (o x o) and DVV (v x v),
     * column-major, as CCDENSC(1) leaves them; the X intermediate as CCDENSZ
     * stores it (-X, a VO vector, index i*nv + a); the Z-vector DVO and DOV
     * (VO vectors) as RPAA leaves them (its totally symmetric block); the MP2
     * amplitudes.  Symmetry-blocked handles: the FOO, FVV, T1 and T2
     * layouts */
    RCC_ARR_DOO = 21, RCC_ARR_DVV = 22, RCC_ARR_XVO = 23, RCC_ARR_DVO = 24,
    RCC_ARR_DOV = 25, RCC_ARR_MP2T1 = 26, RCC_ARR_MP2T2 = 27
};

/* Symmetry description, as RDSYMI / SYMTAB hold it (ccdriv.F:2934, :746).
 * nrep: number of fermion irreps of the abelian double group (1..32);
 * no/nv: occupied / virtual spinors per irrep (NO, NV of symm.inc);
 * frozen: per occupied spinor in irrep-major order, 1 = frozen core (IFROZ,
 * freeze.inc), may be NULL; multb: for nrep > 1, the nrep x nrep pair-irrep
 * table MULTB(1:NREP,1:NREP,1) as RDSYMI leaves it (boson offset removed,
 * values 1..nrep, column-major), may be NULL for nrep = 1.
 * Every array of the interface is then in SYMTAB's symmetry-blocked layout:
 * one block per compound irrep (classes, T2) or per irrep (Fock, T1), blocks
 * consecutive (IOOOOTT .. IVVVVTT, IOO/IVO/IVV, ccdriv.F:1150-1187);
 * rcc_layout_block reports each block's shape and offset.  The orbital
 * energies are in SYMTAB's order: occupied irrep-major, then virtual
 * irrep-major (ccdriv.F:1392-1402). */
typedef struct {
    int nrep;
    const int* no;
    const int* nv;
    const int* frozen;
    const int* multb;
} rcc_sym_t;

/* Layout ids for rcc_layout_block: the six class ids (RCC_OOOO..RCC_VVVV)
 * and the following. */
enum {
    RCC_LAY_T1 = 6,    /* T1, S1, FVO, HOV: NV x NO per irrep (IVO) */
    RCC_LAY_T2 = 7,    /* T2, S2, VVOO: NVVT x NOOT per compound irrep (IVVOOTT) */
    RCC_LAY_FOO = 8,   /* FOO, HO, GO: NO x NO per irrep (IOO) */
    RCC_LAY_FVV = 9    /* FVV, HV, GV: NV x NV per irrep (IVV) */
};

/* Replaces SYMTAB's dimension/offset set-up (ccdriv.F:746-1425) and the
 * CCENER work-array allocation (ccdriv.F:1622-1690) on the device. */
int rcc_create(const rcc_sym_t* sym, rcc_t** handle);
void rcc_destroy(rcc_t* h);
const char* rcc_last_error(const rcc_t* h);

/* MAXDIM (DIIS subspace), MAXIT and NTOL of the &CCENER namelist
 * (ccdriv.F:208-225).  precision: 0 = FP64 everywhere; 1 and 2 run the VVVV
 * ladder, the VOVV.tau product and both (ov)^3 ring contractions as split
 * matrix-core products while denominators, residuals, DIIS and the energy stay
 * FP64.  1 = split-BF16: Re, Im and Re+Im as three BF16 slices each (scale
 * per row and 128-term chunk), FP32 accumulation over 128-term chunks, FP64
 * across chunks.  2 = split-INT8: Re, Im and Re+Im as
 * signed 7-bit digits (rcc_set_split_digits; fixed point per row and 8192-term
 * chunk), exact int32 accumulation, FP64 across chunks.  No counterpart in the
 * reference (FP64 only). */
int rcc_set_options(rcc_t* h, int maxdim, int maxit, int ntol, int precision);

/* How the FP64 complex GEMMs form a complex product: three_mult = 1 (the
 * default) uses Gauss's three real products per complex product (3/4 of the
 * matrix-core work), 0 the classical four.  No counterpart in the reference
 * (its ZGEMM, general/blas.F:13489, is the classical form). */
int rcc_set_complex_product(rcc_t* h, int three_mult);
/* Digits per value of the split-INT8 mode: 4 (28 bits) or 5 (35 bits, the
 * default).  No counterpart in the reference. */
int rcc_set_split_digits(rcc_t* h, int digits);
/* The autotuner of the complex GEMMs (process-wide, on by default unless
 * RCC_AUTOTUNE=0 is in the environment): the first product of each shape
 * met outside graph capture times its candidate work tiles on the stream --
 * the K-split chunk (64 .. 4096) of FP64 products whose output tiles cannot
 * fill the chip, the tile height (32 WM rows, WM 2..4) of split-INT8
 * products -- and keeps the fastest for the process.  Candidates run in
 * place with alpha = 0, beta = 1, which leaves C unchanged.  Different
 * K-splits round differently (FP64 sums in another order).  No counterpart
 * in the reference. */
int rcc_set_autotune(rcc_t* h, int enable);
int rcc_autotune_entries(rcc_t* h, int64_t* fp64_shapes, int64_t* int8_shapes);

/* Fock blocks and orbital energies as FMTOFILE reads them (ccgetv.F:34-87);
 * eps: occupied then virtual (SYMTAB reordering ccdriv.F:1393-1420). */
int rcc_set_fock(rcc_t* h, const double* foo, const double* fvo, const double* fvv,
                 const double* eps);

/* One column batch [col0, col0+ncols) of the block of compound irrep
 * `irrep` (0-based, Fortran IRP - 1) of integral class cls from host memory,
 * as GETVOVV / GETVVVV stream them (ccgetv.F:225-402); a whole block is
 * col0 = 0, ncols = all its columns. */
int rcc_set_ints(rcc_t* h, int cls, int irrep, int64_t col0, int64_t ncols, const double* buf);
/* Same, from device memory already resident on this GPU. */
int rcc_set_ints_device(rcc_t* h, int cls, int irrep, int64_t col0, int64_t ncols,
                        const void* dbuf);

/* Amplitudes in / out (PUTAMPT / GETAMPT ccgetv.F:577-709).  Setting them
 * restarts the iteration from them as CCEQN's entry does (cceqns.F:255-260):
 * iteration count and DIIS history start afresh. */
int rcc_set_amplitudes(rcc_t* h, const double* t1, const double* t2);
int rcc_get_amplitudes(rcc_t* h, double* t1, double* t2);
/* Copy of a device-resident array (RCC_ARR_*) to host memory. */
int rcc_get_array(rcc_t* h, int which, double* out, int64_t nelem_complex);
/* Shape (rows x cols) and offset (complex elements) of block `block`
 * (0..nrep-1) of a layout; block = nrep gives the layout's total size as
 * offset.  Replaces reading SYMTAB's offset arrays (ccdriv.F:1150-1187). */
int rcc_layout_block(rcc_t* h, int lay, int block, int64_t* rows, int64_t* cols,
                     int64_t* offset);

/* MP2EQN (cceqns.F:33-139): T1 = FVO/D, T2 = VVOO/D, ZCORE, MP2ENG energy. */
int rcc_mp2(rcc_t* h, double* emp2);

/* Per-term entry points of one CCEQN iteration (cceqns.F:313-483), acting
 * on the handle's current T1/T2 and its device-resident intermediates. */
int rcc_hointm(rcc_t* h);   /* HOINTM  ccintm.F:455  */
int rcc_hvintm(rcc_t* h);   /* HVINTM  ccintm.F:567  */
int rcc_hovintm(rcc_t* h);  /* HOVINTM ccintm.F:514  */
int rcc_t1eqns(rcc_t* h);   /* T1EQNS  cceqns.F:645 (+T1EQN1 :2521, T1EQN2 :2601) */
int rcc_gointm(rcc_t* h);   /* GOINTM  ccintm.F:95   */
int rcc_gvintm(rcc_t* h);   /* GVINTM  ccintm.F:148  */
int rcc_aintm(rcc_t* h);    /* AINTM   ccintm.F:34   */
int rcc_hintm(rcc_t* h);    /* HINTM   ccintm.F:221  */
int rcc_t2eqns(rcc_t* h);   /* T2EQNS  cceqns.F:937  */
int rcc_bintm(rcc_t* h);    /* BINTM   ccintm.F:625  */
int rcc_denom(rcc_t* h);    /* DENOM   cceqns.F:2355 */
int rcc_zcore(rcc_t* h);    /* ZCORE   cceqns.F:1864 */
int rcc_diis(rcc_t* h, double* rms);       /* DIISX  cceqns.F:1985 */
int rcc_energy(rcc_t* h, double* ecc);     /* CCENG  cceqns.F:2180 (on S) */
int rcc_t1diag(rcc_t* h, double* diag);    /* T1DIAG cceqns.F:2453 (on S) */
/* PUT S1,S2 INTO T1,T2 (cceqns.F:560-561). */
int rcc_accept(rcc_t* h);

/* One whole iteration (HOINTM .. T1DIAG, then S -> T) = one pass of the
 * CCEQN loop body (cceqns.F:293-569).  nit counts calls since rcc_mp2. */
int rcc_iterate(rcc_t* h, int* nit, double* ecc, double* rms);

/* Replay the term sequence of rcc_iterate / rcc_cceqn / rcc_iterate_batch
 * (HOINTM .. BINTM, DENOM: all device work) from a hipGraph: the first
 * iteration runs eagerly, the second is captured, later ones are one graph
 * launch.  Any change of integrals or precision options re-captures.  No
 * counterpart in the reference. */
int rcc_set_graph(rcc_t* h, int enable);
/* One CCEQN iteration of each of n independent handles (each on its own
 * stream): every handle's term sequence is enqueued first, then DIISX,
 * CCENG and S -> T run handle by handle while the other handles' terms are
 * still on the device.  nit/ecc/rms receive one value per handle.  No
 * counterpart in the reference (which runs one system per process). */
int rcc_iterate_batch(rcc_t* const* handles, int n, int* nit, double* ecc, double* rms);
/* As rcc_iterate_batch, with the four dominant products of every handle
 * (VOVV.tau, the HINTM and T2EQNS rings, the VVVV ladder; per irrep block
 * for a symmetric handle) run as one grouped launch per term over all the
 * handles.  Handles must use precision 0 and the same complex-product form
 * and, with symmetry, self-inverse irreps.  The first call per list of
 * handles builds the launch descriptors.  Same results as rcc_iterate. */
int rcc_iterate_batch_grouped(rcc_t* const* handles, int n, int* nit, double* ecc, double* rms);
/* The plan rcc_iterate_batch_grouped built for these handles: info[0] = 1 if
 * the term sequence was compiled (else the segmented form runs), info[1] =
 * launch positions of the compiled term sequence (each one launch for all
 * handles), info[2] = grouped products among them, info[3..5] = how many of
 * those the tile autotuner put on the 96 x 64 tile, the 32 x 32 tile and the
 * persistent 32 x 32 grid, info[6] = how many run the streaming kernel
 * (products with an occupied-spinor block width; fixed, not tuned), info[7] =
 * how many the autotuner put on the 128 x 32 tile.  info holds 8 ints. */
int rcc_batch_plan_info(rcc_t* const* handles, int n, int* info);
/* CCEQN's loop (cceqns.F:293-569): iterate until rms < 10^-ntol or maxit. */
int rcc_cceqn(rcc_t* h, int* nit, double* ecc, double* rms);

/* ---- CCSD Lambda equations and the one-particle response density.
 * relccsd holds Lambda branches (CCEQN cceqns.F:277-288, T1EQNS :860-932,
 * the Lambda-only routines of ccintm.F) but switches them off
 * (ccdriv.F:1918-1924); these entry points have no live counterpart.  They
 * solve the stationarity conditions of the CCSD Lagrangian
 *     L = E(T) + Lam . (S(T) - D o T),
 * S the pre-DENOM right-hand side of one CCEQN pass, D the DENOM
 * denominators, E CCENG's complex sum:  Lam = (dE/dT + Lam . dS/dT) / D,
 * iterated with DENOM, ZCORE and DIISX; Lam . dS/dT is the adjoint sweep of
 * the iteration's own contractions at the handle's T (with symmetry, the
 * ladder's adjoint per compound irrep block in one grouped launch).
 * Precision 0.  Lam is in the T1 / T2 layouts (symmetry-blocked for
 * nrep > 1). */
/* Solve for Lam at the current T (normally after rcc_cceqn) with the
 * handle's MAXIT / MAXDIM / NTOL; nit and rms of the last step.  Clears the
 * DIIS history (a later rcc_cceqn restarts it). */
int rcc_lambda(rcc_t* h, int* nit, double* rms);
int rcc_set_lambda(rcc_t* h, const double* l1, const double* l2);
int rcc_get_lambda(rcc_t* h, double* l1, double* l2);
/* One adjoint sweep at the current T for the given Lam (no DENOM):
 * g = with_energy * dE/dT + Lam . dS/dT. */
int rcc_lambda_vjp(rcc_t* h, const double* l1, const double* l2, int with_energy, double* g1,
                   double* g2);
/* One-particle response density of the current (T, Lam): dens is n x n
 * complex (n = no + nv, occupied then virtual in the order of the orbital
 * energies of rcc_set_fock, column-major), Hermitian, with
 * d Re(Ec) / d lam = tr(dens V) for the one-electron perturbation
 * f -> f + lam V (FOO, FVO, FVV and the orbital energies shifted by the
 * blocks and the diagonal of the Hermitian V). */
int rcc_density(rcc_t* h, double* dens);
/* Device time of the last rcc_lambda iteration, ms. */
int rcc_last_lambda_ms(rcc_t* h, double* ms);

/* ---- Integral front end: RDINTS (ccints.F:33-569) with MKINDEX (:1421),
 * GET2INTS (:765) and GETFM / MAKEFM (:1031-1419), called by CCSETI
 * (ccdriv.F:1427-1570) when DOSORT.  Reads the MOLFDIR files MRCONEE
 * (records 1-6: dimensions, irreps and orbital energies of the spinors,
 * the one-electron matrix) and MDCINT (the Kramers-reduced two-electron
 * list), expands every canonical integral into its 16 time-reversal /
 * permutation / conjugation images on the GPU, antisymmetrises them into
 * the six classes of the handle (replacing what rcc_set_ints would set),
 * and builds FOO, FVO, FVV and the orbital energies (Re diag F) as GETFM
 * does with USEOE = NORECMP = F.  nelec[nnelec]: occupied spinors per
 * irrep of MRCONEE record 2 (the &RELCCSD NELEC input); the handle's no / nv
 * must match.  e0e1 (may be null, 3 doubles): E0, E1 and ECORE of RDINTS's
 * summary (ESCF = E0 + E1 + ECORE).  When two images land on one element,
 * the later one in file order is kept, as relccsd's sequential stores do.
 * For a symmetry-blocked handle the files' abelian irreps (MRCONEE records
 * 3-5) must give the handle's occupied / virtual counts per irrep. */
int rcc_read_molfdir(rcc_t* h, const char* mrconee, const char* mdcint, const int* nelec,
                     int nnelec, double* e0e1);
/* rcc_read_molfdir with namelist CCSORT's options (ccdriv.F:155, 342-343;
 * GETFM ccints.F:1080, 1150-1198, 1260, 1287): useoe keeps MRCONEE's orbital
 * energies instead of Re diag F; norecmp takes F = diag of those energies
 * (FVO = 0) and E0 = their occupied sum instead of rebuilding F (E1 still
 * from the integrals).  rcc_read_molfdir is this with both 0. */
int rcc_read_molfdir_opts(rcc_t* h, const char* mrconee, const char* mdcint, const int* nelec,
                          int nnelec, int useoe, int norecmp, double* e0e1);
/* The expansion and sort alone, from a canonical list already in device
 * memory: d_idx m x (IKR, JKR, KKR, LKR) int32, d_val m complex; kr the
 * MDCINT header's (KR(i), KR(-i)) pairs (1-based spinors), spinor_pos[s]
 * the handle position of spinor s (>= 0 occupied, -(virtual position + 1)).
 * ms (may be null): device time of the sort. */
int rcc_sort_ints_device(rcc_t* h, const int* spinor_pos, int nsp, const int* kr, int nkr,
                         const void* d_idx, const void* d_val, int64_t m, double* ms);
/* Device time (ms) of the last integral sort. */
double rcc_last_sort_ms(rcc_t* h);
/* ---- One-electron properties.  rcc_read_property: RDPROP (ccints.F:569-764)
 * -- operator `name` from an MDPROP file, returned as the (no + nv) x (no + nv)
 * complex matrix in the handle's spinor order (occupied, then virtual;
 * column-major), with the spinor classification of MRCONEE + nelec as in
 * rcc_read_molfdir.  rcc_expectation: EXPVAL (ccgrad.F:261-347) -- level 0,
 * the reference-determinant value sum_i P(i,i) (CCGRAD's SCF line,
 * ccdriv.F:2028-2039); level 1, the unrelaxed CCSD value sum_i P(i,i) +
 * tr(D P) with D the Lambda response density of rcc_density (needs
 * rcc_lambda first), i.e. d(E_ref + Ec)/d lam for f -> f + lam P at fixed
 * orbitals (no orbital relaxation); level 2, the MP2 relaxed value of
 * relccsd's CCFOPR (ccdriv.F:2080-2087), after rcc_mp2_density.
 * value: 2 doubles (Re, Im). */
int rcc_read_property(rcc_t* h, const char* mrconee, const char* mdprop, const char* name,
                      const int* nelec, int nnelec, double* prop);
int rcc_expectation(rcc_t* h, const double* prop, int level, double* value);
/* ---- MP2 relaxed density: the DOMP2G block of CCFOPR (ccdriv.F:2058-2090),
 * replacing MP2EQN with EQNS = 'LAMBDA' (cceqns.F:33), CCDENSC(1)
 * (ccgrad.F:34) and CCDENSZ(1) (ccgrad.F:141) with its Z-vector solve RPAA
 * (ccrpa.F:34) at zero frequency, on the handle's integrals and orbital
 * energies (its amplitudes are not touched); C1 or symmetry-blocked.  maxit / maxdim / ntol
 * are RPAA's MXITGR / MXDIMGR / NTOLGR (converged when the DIIS RMS drops
 * below 10^-ntol; maxdim <= 16).  nit, rms (may be null): iterations taken,
 * last RMS.  Afterwards rcc_expectation level 2 gives the MP2 relaxed value
 * of EXPVAL, and rcc_get_array RCC_ARR_DOO .. RCC_ARR_MP2T2 the pieces. */
int rcc_mp2_density(rcc_t* h, int maxit, int maxdim, int ntol, int* nit, double* rms);
/* Device time (ms) of the last rcc_mp2_density; rcc_mp2d_timing: 6 doubles,
 * total, MP2 amplitudes + densities, X intermediate, Z-vector loop, the sum
 * over the Z-vector's matrix-vector launches, and the VOVV x lambda GEMM
 * (part of the X intermediate). */
double rcc_last_mp2d_ms(rcc_t* h);
int rcc_mp2d_timing(rcc_t* h, double* ms);
/* The handle's orbital energies (no + nv doubles, occupied then virtual). */
int rcc_get_eps(rcc_t* h, double* eps);

/* ---- MRCONEE writer (host only, no GPU): records 1-6 as RDSYMI
 * (ccdriv.F:2978-2988) and GETFM (ccints.F:1091-1096) read them, i.e. the
 * file tmoone's MFDTRAN writes for relccsd (tmoone/mfdtran.F:219-226): NSP,
 * BREIT, ECORE; the irrep names and occupied count per irrep (nelec, nrep
 * ints); the abelian irrep names; MULTB(1:2NREP, 1:2NREP, 1) (column-major;
 * null: the C1 table, nrep = 1); per spinor its irrep (1..nrep, also as the
 * higher-group irrep) and orbital energy; the NSP x NSP one-electron matrix
 * h1 (complex), record 6 holding h1(p,q) at p + NSP q, which is how GETFM
 * reads it (MFDTRAN writes ONE(i,j) with j fastest, so h1 = ONE^T gives its
 * file).  Errors through rcc_file_error(). */
int rcc_mrconee_write(const char* path, int nsp, int nrep, const int* irreps, const int* nelec,
                      const int* multb, const double* eps, const double* h1, double ecore,
                      int breit);

/* ---- MCCRES restart file: PUTAMPT / GETAMPT (ccgetv.F:577-709) and the
 * status / energy records of CCDUMP (ccdriv.F:2767-2816).  Direct access,
 * 8192-byte records (RSRCLEN = 1024 REAL*8, IRECLEN for LINUX): record 1
 * unused, 2 IMSTAT(1:30) (int32), 3 ETOT ESCF EMP2 ECCSD ET1 ET2 ET3, from 4
 * T1 and from 4 + (2 NDIMT1 / 1024 + 1) T2, 512 complex per record, in
 * relccsd's (symmetry-blocked) amplitude layout.  The handle forms read /
 * write the handle's current T1, T2; imstat (30 ints) and energies (7
 * doubles) may be null (written as zeros).  The host forms need no GPU and
 * report errors through rcc_file_error(). */
int rcc_write_mccres(rcc_t* h, const char* path, const int* imstat, const double* energies);
int rcc_read_mccres(rcc_t* h, const char* path, int* imstat, double* energies);
int rcc_mccres_write(const char* path, int64_t ndimt1, int64_t ndimt2, const double* t1,
                     const double* t2, const int* imstat, const double* energies);
int rcc_mccres_read(const char* path, int64_t ndimt1, int64_t ndimt2, double* t1, double* t2,
                    int* imstat, double* energies);
const char* rcc_file_error(void);

/* ---- Non-iterative triples: T3CORR (cceqns.F:1336-1863), called by
 * CCENER after the CCSD equations when DOCCSDT is set (ccdriv.F:1799-1837).
 * On the handle's current T1 / T2 (normally after rcc_cceqn), the set Fock
 * FVO, orbital energies and the VOOO, VVOO, VOVV classes:
 *   et[0] = ET1, the fourth-order +T correction     ('4th order triples')
 *   et[1] = ET2, the fifth-order (T) term          (CCSD(T) = Ec + ET1 + ET2)
 *   et[2] = ET3, the fifth-order -T term           (CCSD-T = Ec + ET1 + ET3)
 * as T3CORR writes them to COMMON /ENERG/ (cceqns.F:1799-1812).  FP64 on the
 * matrix cores for every precision mode; symmetric handles run the C1 sums
 * (symmetry-forbidden elements are zero).  Occupied triples i > j > k are
 * numbered in the reference's loop order (k outer, then j, then i);
 * rcc_t3corr_part sums the range [first, first + count) only (count < 0:
 * to the end), so a set of processes can split the triples and add their
 * three numbers.  Returns 0 and zeros for fewer than 3 occupied or 3
 * virtual spinors. */
int rcc_t3corr(rcc_t* h, double* et);
int rcc_t3corr_part(rcc_t* h, int64_t first, int64_t count, double* et);
/* Number of occupied triples (o (o-1) (o-2) / 6). */
int64_t rcc_t3_triples(rcc_t* h);
/* Device time (ms) of the last rcc_t3corr call; parts (may be null, 4
 * doubles): [0] its triples GEMMs, [1] its energy kernels (with their
 * reductions), [2] occupied triples per launch, [3] launches. */
double rcc_last_t3_ms(rcc_t* h, double* parts);

/* Device timing of the last rcc_iterate / rcc_cceqn iteration, ms (after a
 * grouped batch iteration: that batch iteration, the same for every handle). */
int rcc_last_iteration_ms(rcc_t* h, double* ms);
/* HIP-event durations (ms, on the handle's stream) of the last iteration's
 * dominant launches: [0] VVVV ladder (BINTM), [1] HINTM (ov)^3 ring,
 * [2] T2EQNS (ov)^3 ring, [3] BINTM VOVV.tau product. */
int rcc_kernel_times(rcc_t* h, double* ms4);

/* Diagnostics: the library's complex GEMM engine on device pointers
 * (column-major, op 0=N 1=T 2=C 3=conjugate without transpose, 3 only with 0
 * on the other operand), and the MFMA lane-layout probe. */
int rcc_zgemm_device(rcc_t* h, int opa, int opb, int m, int n, int k, const double* alpha,
                     const void* A, int64_t lda, const void* B, int64_t ldb,
                     const double* beta, void* C, int64_t ldc);
/* Allocation-free form of rcc_zgemm_device.  No counterpart in relccsd: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Same product and conventions; it calls no
 * hipMalloc / hipFree and does not synchronise: the product is enqueued on
 * the handle's stream (rcc_synchronize waits for it).  The K-split partial
 * sums go to a fixed 32 MiB scratch array per device, shared by every call
 * on that device: calls that overlap on one GPU race.  ksplit = number of k
 * chunks (0: the library's heuristic split that fits the scratch); a split
 * whose partial sums (ksplit x m x n x 16 bytes) exceed 32 MiB is refused
 * with an error.  No autotuning. */
int rcc_zgemm_device_noalloc(rcc_t* h, int opa, int opb, int m, int n, int k,
                             const double* alpha, const void* A, int64_t lda, const void* B,
                             int64_t ldb, const double* beta, void* C, int64_t ldc, int ksplit);
/* The split-BF16 GEMM on device pointers: C = alpha A B + beta C with A (m x k)
 * and B (k x n) column-major complex (diagnostic; allocates its own planes and
 * room for an 8-way K-split's partial sums). */
int rcc_zgemm_split_device(rcc_t* h, int m, int n, int k, const double* alpha, const void* A,
                           int64_t lda, const void* B, int64_t ldb, const double* beta, void* C,
                           int64_t ldc);
/* The split-INT8 GEMM on device pointers, same conventions (diagnostic). */
int rcc_zgemm_i8_device(rcc_t* h, int digits, int m, int n, int k, const double* alpha,
                        const void* A, int64_t lda, const void* B, int64_t ldb,
                        const double* beta, void* C, int64_t ldc);
int rcc_probe_mfma(rcc_t* h, double* out256);
/* Achievable FP64 matrix-core rate: nblocks x 4 waves, each running iters
 * rounds of `chains` independent v_mfma_f64_16x16x4_f64 (1, 2, 4, 8, 12). */
int rcc_probe_mfma_rate(rcc_t* h, int chains, int iters, int nblocks, double* tflops);
int rcc_synchronize(rcc_t* h);

#ifdef __cplusplus
}
#endif

#endif /* RELCCSD_HIP_H */
