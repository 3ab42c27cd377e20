// Internal declarations shared by the relccsd HIP library sources.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

namespace rcc {

// op(X): N as stored, T transposed, C conjugate-transposed, R conjugated (not
// transposed; the adjoint contractions of the Lambda equations use it)
enum { OP_N = 0, OP_T = 1, OP_C = 2, OP_R = 3 };

// packed index of the pair x > y of 0 .. n-1 (y-major triangle: the relccsd
// A>B / I>J packing of the T2 and VVVV dimensions)
__host__ __device__ __forceinline__ int64_t pair_of(int x, int y, int n) {
    return (int64_t)y * (2 * n - y - 1) / 2 + (x - y - 1);
}

// Strided-batched complex GEMM, column-major: C_b = alpha op(A_b) op(B_b) + beta C_b.
struct GemmArgs {
    int opA = OP_N, opB = OP_N;
    int m = 0, n = 0, k = 0;
    const double2* A = nullptr; int64_t lda = 0, sA = 0;
    const double2* B = nullptr; int64_t ldb = 0, sB = 0;
    double2* C = nullptr; int64_t ldc = 0, sC = 0;
    double2 alpha = {1.0, 0.0}, beta = {0.0, 0.0};
    int batch = 1;
    // distinct kernel symbol for profiling: 0 generic, 1 VVVV ladder, 2 (ov)^3 rings
    int tag = 0;
    // 1: Gauss three-multiplication complex product (3/4 of the MFMA work)
    int gauss = 0;
};

// ws may be null (no K-split); ws_elems = capacity in complex elements.
int zgemm(const GemmArgs& g, double2* ws, int64_t ws_elems, hipStream_t st);
// The allocation-free form of zgemm: K-split partial sums in a fixed
// file-scope device array of zgemm_scratch_elems() complex elements per
// device.  zgemm_scratch_init(device) (device current) takes the array's
// address once per device; zgemm_noalloc then splits k into ksplit chunks
// (0: the heuristic split that fits the array) and fails with err set when
// the partial sums would not fit.  No allocation, no autotuning.
int zgemm_scratch_init(int device);
int64_t zgemm_scratch_elems();
// Workgroups of the FP64 GEMM engine resident at once on the current device
// (two per CU).
int64_t resident_slots();
int zgemm_noalloc(const GemmArgs& g, int device, int ksplit, hipStream_t st, std::string& err);
// K-split autotuner of zgemm (process-wide): on / off, and the number of
// shapes tuned so far.
void zgemm_autotune(int enable);
bool zgemm_autotune_enabled();
int64_t zgemm_tuned_shapes();

// Grouped GEMM: many independent C_g = alpha_g A_g B_g + beta_g C_g (op(A) = N,
// op(B) = N or C) in one launch.  Group g owns blocks [tile_off, tile_off +
// tiles(g)) of the grid, tile_off a multiple of 8 so the XCD-aware tile order
// holds inside each group.
struct ZGroup {
    int m, n, k, pad;
    int64_t lda, ldb, ldc;
    const double2* A;
    const double2* B;
    double2* C;
    double2 alpha, beta;
    int64_t tile_off;
};
// Sets tile_off of every group; returns the grid size (blocks).
int64_t zgroup_offsets(ZGroup* g, int ngroups);
// dG: device copy of the descriptors (with offsets); tag as GemmArgs::tag.
// gmap (optional): the group of every block, instead of a binary search.
int zgemm_grouped(int opb, int tag, int gauss, const ZGroup* dG, int ngroups,
                  int64_t blocks, hipStream_t st, const int* gmap = nullptr);
// Grouped GEMM with op(A) and op(B) given (N/T/C/R combinations the term
// sequence uses: NN, NC, NT, TN, CN); tag as GemmArgs::tag for NN / NC.
int zgemm_grouped_ops(int opa, int opb, int tag, int gauss, const ZGroup* dG, int ngroups,
                      int64_t blocks, hipStream_t st, const int* gmap = nullptr);
// The 32 x 32-tile grouped GEMM (same op combinations), its tile offsets;
// persistent > 0: a grid of that many workgroups striding over the tiles.
int64_t zgroup_offsets_small(ZGroup* g, int ngroups);
int zgemm_grouped_small(int opa, int opb, int gauss, const ZGroup* dG, int ngroups,
                        int64_t blocks, hipStream_t st, const int* gmap, int persistent = 0);
// The 128 x 32-tile grouped GEMM (same op combinations and bits as the
// 96 x 64 one) and its tile offsets.
int64_t zgroup_offsets_tall(ZGroup* g, int ngroups);
int zgemm_grouped_tall(int opa, int opb, int gauss, const ZGroup* dG, int ngroups,
                       int64_t blocks, hipStream_t st);
// Streaming grouped product (op N N, n <= nj <= 16, one row per thread,
// FP64 vector FMAs), its offsets in 256-row blocks.
int64_t zgroup_offsets_skinny(ZGroup* g, int ngroups);
int zgemm_grouped_skinny(int nj, const ZGroup* dG, int ngroups, int64_t blocks, hipStream_t st);
// ---- non-iterative triples (rcc_triples.hip) ----
// Inputs of T3CORR in the C1 layouts (device pointers); gauss as GemmArgs.
struct T3Input {
    int no = 0, nv = 0, gauss = 1;
    const double2 *T1 = nullptr, *T2 = nullptr, *FVO = nullptr;
    const double2 *VOOO = nullptr, *VVOO = nullptr, *VOVV = nullptr;
    const double* eps = nullptr;
};
struct T3Timing {
    double gemm_ms = 0, energy_ms = 0, total_ms = 0;   // GEMMs, energy + reduce, whole call
    int64_t chunks = 0;
    int chunk = 0;                                      // occupied triples per launch
};
struct T3State;
T3State* t3_new();
void t3_free(T3State* s);
int64_t t3_triples(int no);
int t3_chunk(const T3State* s);
// ET1, ET2, ET3 of the occupied triples [first, first + count) (reference
// loop order); work buffers are kept in s between calls.
int t3_run(T3State* s, const T3Input& in, int64_t first, int64_t count, hipStream_t st,
           double* et, T3Timing* tm, std::string& err);
// ---- integral front end (rcc_frontend.hip) ----
struct MolfdirOne {          // MRCONEE records 1-6
    int nsp = 0, nsymrp = 0, nrep = 0;
    bool breit = false;
    double ecore = 0.0;
    std::vector<int> irrep_higher, irrep;   // 1-based, per spinor
    std::vector<double> eps;
    std::vector<double2> h1;                // nsp x nsp, column-major
};
struct MdcintList {          // MDCINT: KR table and the canonical integrals
    int nkr = 0;
    std::vector<int> kr;                    // (KR(i), KR(-i)), i = 1..nkr, 1-based spinors
    std::vector<int4> idx;                  // (IKR, JKR, KKR, LKR)
    std::vector<double2> val;
};
int read_mrconee(const char* path, MolfdirOne& out, std::string& err);
// header_only_ok: a header record without the KR table is accepted (MDBINT)
int read_mdcint(const char* path, MdcintList& out, std::string& err,
                bool header_only_ok = false);
// RDPROP: the NSP x NSP matrix of operator `name` on an MDPROP file
int read_mdprop(const char* path, const char* name, int nsp, std::vector<double2>& prop,
                std::string& err);
// Expansion + sort of m canonical integrals (device arrays) into the six
// packed C1 classes cls[0..5] (OOOO, VOOO, VVOO, VOVO, VOVV, VVVV), which
// are overwritten.  d_spos[s]: C1 position of spinor s (>= 0 occupied,
// -(position + 1) virtual); d_kr[nkr + i] = spinor of KR(i), i in [-nkr, nkr].
// itrsgn: GET2INTS's sign of the time-reversed pair images (1 Coulomb list,
// -1 Breit list).
int sort_integrals_device(int no, int nv, double2* const* cls, const int* d_spos, int nsp,
                          const int* d_kr, int nkr, const int4* d_idx, const double2* d_val,
                          int64_t m, hipStream_t st, double* ms, std::string& err,
                          int itrsgn = 1);
// GETFM + MAKEFM on the sorted classes: Fock blocks, orbital energies, E0, E1.
int build_fock_device(int no, int nv, int nsp, const double2* d_h1, const int* d_occ,
                      const int* d_vir, double2* const* cls, double2* FOO, double2* FVO,
                      double2* FVV, double* eps, double* d_e01, hipStream_t st, std::string& err);
int mfma_probe(double* dout, hipStream_t st);
int mfma_rate(int chains, int iters, int nblocks, double* out, hipStream_t st);

// ---- launch recording and batched replay (rcc_kernels.hip) ----
// With g_record set (per host thread), the element-wise wrappers below,
// zgemm and the solver's grouped launches append a Rec instead of launching.
// The recordings of several handles' term sequences (identical in kind,
// position by position) compile into a BatchProgram that issues each
// position once for all handles.
enum RecKind {
    RK_GETTAU, RK_EXPAND_E2, RK_EXPAND_F1, RK_EXPAND_RING, RK_PACK_PIJ, RK_PACK_PAB,
    RK_PACK_PABIJ, RK_PERM_AKCI, RK_TRACE_T1, RK_TRACE_GV, RK_DENOM1, RK_DENOM2, RK_AXPBY,
    RK_CONJ, RK_COPY, RK_ZERO_DIAG, RK_BLOCKSET, RK_FILL, RK_SPLIT_REDUCE, RK_GEMM, RK_GROUPED,
    // batched DIISX / CCENG stages (not recorded; parameters built per iteration)
    RK_ZDOT_PART, RK_FINISH, RK_SUMNORM_PART, RK_LINCOMB, RK_DIIS_PREP,
    // the symmetry-blocked engine's element-wise work (rcc_blocked.hip)
    RK_MAPSUM, RK_MAPBIL, RK_BDENOM, RK_GRED
};
struct Rec {
    int kind = 0;
    int64_t n = 0;                    // element-wise: elements (grid x extent)
    int gy = 1;                       // element-wise: grid y of the kernel
    std::vector<unsigned char> blob;  // element-wise: kernel parameters
    GemmArgs g;                       // RK_GEMM: the product, its K split and workspace
    int nsplit = 1, kchunk = 0;
    double2* ws = nullptr;
    std::vector<ZGroup> groups;       // RK_GROUPED: descriptors, op(A), op(B), tag, complex form
    int opa = 0, opb = 0, tag = 0, gauss = 0;
    // RK_GROUPED: -1 the batch tuner picks the tile (the MFMA tiles give the
    // same bits); 3 the streaming kernel (vector FMAs: fixed, nj columns)
    int variant = -1, nj = 0;
};
extern thread_local std::vector<Rec>* g_record;
struct BatchProgram;
int batch_compile(const std::vector<std::vector<Rec>*>& recs, BatchProgram** out,
                  std::string& err);
int batch_run(const BatchProgram* p, hipStream_t st);
// times the 96 x 64 and 32 x 32 tiles of every grouped product, keeps the faster
int batch_tune(BatchProgram* p, hipStream_t st);
// positions (launch sites) of the program, grouped products among them, and
// how many of those use each kernel (96 x 64, 32 x 32, 32 x 32 persistent,
// streaming)
void batch_info(const BatchProgram* p, int* positions, int* grouped, int* variants);
void batch_free(BatchProgram* p);

// ---- element-wise / permutation kernels (rcc_kernels.hip) ----
// Packed pair tables: pair p of an n-set -> (x, y), x > y, y outer.
struct PairTab {
    const int2* tab;   // device, npair entries
    int n;
    int64_t npair;
};

// Parameter blocks of the reductions and updates behind DIISX and CCENG, for
// launching them over many handles at once (k_many).  Sizes by kind:
// batch_param_size.  The single-handle k_* functions below use the same
// kernels with the same partitioning, so batched and single results agree
// bit for bit.
int red_blocks(int64_t n, int nblk);
void zdot_params(const double2* const* xs, int nx, const double2* y, int64_t n, double2* dst,
                 double2* part, int nblk, void* zdot_part, void* finish);
void sumnorm_params(const double2* x, int64_t n, double2* dst, double2* part, int nblk,
                    void* sumnorm_part, void* finish);
void lincomb_params(const double2* const* xs, const double* coef, int nx, double2* y, int64_t n,
                    void* lincomb);
void diis_prep_params(const double2* S, const double2* T, double2* d, double2* amp, int64_t n,
                      void* prep);
void gettau_params(const double2* t1, const double2* t2, double2* tau, double s2, PairTab vv,
                   PairTab oo, int nv, void* gettau);
void copy_params(int64_t n, const double2* x, double2* y, void* copy);
size_t batch_param_size(int kind);
// gx: largest n (grid-stride kinds) or largest block count (reductions)
int k_many(int kind, const void* dparams, int count, int64_t gx, hipStream_t st);


void k_gettau(const double2* t1, const double2* t2, double2* tau, double s2,
              PairTab vv, PairTab oo, int nv, hipStream_t st);
// out[r + R*(y + no*x)] = sgn * P[r, pair(x,y)]  (0 for x == y)
void k_expand_e2(const double2* P, double2* out, int64_t R, PairTab oo, int no, hipStream_t st);
// out[a + nv*(d + nv*col)] = sgn * P[pair(a,d), col]
void k_expand_f1(const double2* P, double2* out, int64_t ncol, PairTab vv, int nv,
                 hipStream_t st);
// out[(a + nv*i) + nvo*(c + nv*k)] = s * t2[a,c,i,k] - (t1 ? t1[a,k] t1[c,i] : 0)
void k_expand_ring(const double2* T2, const double2* t1, double s, double2* out, PairTab vv,
                   PairTab oo, int nv, int no, hipStream_t st);
// out[row + R*(c + nv*d)] = sgn * P[row, pair(c,d)]   (column pair expansion)
void k_expand_cols(const double2* P, double2* out, int64_t R, PairTab vv, int nv,
                   hipStream_t st);
// S[R, pair(i,j)] += alpha * (Y[R + NR*(i + no*j)] - Y[R + NR*(j + no*i)])
void k_pack_pij(const double2* Y, double2* S, int64_t NR, double alpha, PairTab oo, int no,
                hipStream_t st);
// S[pair(a,b), col] += alpha * (Y[a + nv*(b + nv*col)] - Y[b + nv*(a + nv*col)])
void k_pack_pab(const double2* Y, double2* S, int64_t ncol, double alpha, PairTab vv, int nv,
                hipStream_t st);
// S[ab, ij] += alpha (Q[ai,bj] - Q[bi,aj] - Q[aj,bi] + Q[bj,ai])   (ring layout Q)
void k_pack_pabij(const double2* Q, double2* S, PairTab vv, PairTab oo, int nv, int no,
                  hipStream_t st, double alpha = 1.0);
// H[(a,i),(c,k)] (+)= X[(a,k),(c,i)]
void k_perm_akci(const double2* X, double2* H, int nv, int no, bool accumulate,
                 hipStream_t st);
// S1[a,i] += sum_k sgn X[(a,k), pair(i,k)]
void k_trace_t1eqn2(const double2* X, double2* S1, PairTab oo, int nv, int no, hipStream_t st);
// GV[a,c] += alpha * sum_k H1[(a,k),(c,k)]   (H1 in (AK,CI) layout)
void k_trace_gv(const double2* H1, double2* GV, double alpha, int nv, int no, hipStream_t st);
// transposes (adjoints) of k_trace_t1eqn2 and k_trace_gv:
// Xb[(a,y), pair(x,y)] += S1b[a,x] and Xb[(a,x), pair(x,y)] -= S1b[a,y]  (x > y)
void k_trace_t1eqn2_adj(const double2* S1b, double2* Xb, PairTab oo, int nv, int no,
                        hipStream_t st);
// H1b[(a,k),(c,k)] += alpha * GVb[a,c]
void k_trace_gv_adj(const double2* GVb, double2* H1b, double alpha, int nv, int no,
                    hipStream_t st);
// z = x * y element-wise (complex)
void k_hadamard(int64_t n, const double2* x, const double2* y, double2* z, hipStream_t st);
// eps == nullptr: only the frozen-spinor zeroing (ZCORE); frozen == nullptr: none
void k_denom(double2* S1, double2* S2, const double* eps, const int* frozen, PairTab vv,
             PairTab oo, int nv, int no, hipStream_t st);
// y = a*x + b*y (complex a, b)
void k_axpby(int64_t n, double2 a, const double2* x, double2 b, double2* y, hipStream_t st);
void k_conj(int64_t n, const double2* x, double2* y, hipStream_t st);
// y = x (device to device, as a kernel so that it records and replays)
void k_copy(int64_t n, const double2* x, double2* y, hipStream_t st);
// M (n x n) := M with the diagonal zeroed (square, column-major)
void k_zero_diag(double2* M, int n, hipStream_t st);
// generic gather: out[i] = sgn[i] * (conj ? conj(in[idx[i]]) : in[idx[i]]), idx<0 -> 0
void k_gather(const double2* in, const int64_t* idx, const signed char* sgn, bool conj,
              double2* out, int64_t n, hipStream_t st);
// part[b] = sum over chunk b of conj(x) * y   (deterministic, fixed order)
void k_zdotc_multi(const double2* const* xs, int nx, const double2* y, int64_t n,
                   double2* dst, double2* part, int nblk, hipStream_t st);
void k_sum_and_norm(const double2* x, int64_t n, double2* dst, double2* part, int nblk,
                    hipStream_t st);
void k_lincomb(const double2* const* xs, const double* coef, int nx, double2* y, int64_t n,
               hipStream_t st);

// ---- symmetry-blocked engine kernels ----
// out[e] = beta out[e] + alpha sum_t f_t(src[slot[t]][map[t n + e] >> 2]), map
// entries (index << 2) | (negate << 1) | conjugate, -1 for a zero term,
// term-major; T > 4 terms all read src[0].
void k_map_sum(double2* out, const double2* const* src, const int* slot, int nsrc,
               const int32_t* map, int64_t n, int T, double alpha, double beta, hipStream_t st);
// out[e] = c0 f(src0[m0]) + sg0 A[p0] B[q0] + sg1 A[p1] B[q1] (np terms, -1: absent)
void k_map_bil(double2* out, const double2* src0, const double2* A, const double2* B,
               const int32_t* m0, const int32_t* p0, const int32_t* q0, const int32_t* p1,
               const int32_t* q1, int64_t n, int np, double c0, double sg0, double sg1,
               hipStream_t st);
void mapbil_params(double2* out, const double2* src0, const double2* A, const double2* B,
                   const int32_t* m0, const int32_t* p0, const int32_t* q0, const int32_t* p1,
                   const int32_t* q1, int64_t n, int np, double c0, double sg0, double sg1,
                   void* q);
// S[e] /= D[e] (divide), S[e] = 0 where F[e] != 0 (mask)
void k_bdenom(double2* S, const double* D, const double* F, int64_t n, int divide, int mask,
              hipStream_t st);
// K-split partial sums of grouped products: C = alpha sum_s W[s m n + e] + beta C
struct GRedDesc {
    double2* C;
    const double2* W;
    int64_t ldc;
    int m, n, nsplit, pad;
    double alpha, beta;
};
// maxn: the largest m n split_ways(nsplit) over the descriptors
void k_gred(const GRedDesc* d, int count, int64_t maxn, hipStream_t st);

// symmetry blocks (packed, column-major, rows x ncols) <-> C1 arrays:
// C1 element of block element (r, c) = rowmap[r] + ld * colmap[c]
void k_scatter_block(const double2* src, int64_t rows, int64_t ncols, const int* rowmap,
                     const int* colmap, double2* dst, int64_t ld, hipStream_t st);
void k_gather_block(const double2* src, int64_t ld, const int* rowmap, const int* colmap,
                    int64_t rows, int64_t ncols, double2* dst, hipStream_t st);
void k_fill_zero(double2* p, int64_t n, hipStream_t st);

// MP2 relaxed property (rcc_props.hip).  X[a + nv*i] += sum_kl <ak||il> D[k + no*l]
// over the packed VOOO class (part: 2 noot nv scratch); X[a + nv*i] += sum_bc
// D[b + nv*c] conj <bi||ca> over the packed VOVV class (part: 2 nvvt no scratch)
void k_x_vooo_doo(const double2* vooo, const double2* doo, double2* x, PairTab oo, double2* part,
                  int nv, int no, hipStream_t st);
void k_x_vovv_dvv(const double2* vovv, const double2* dvv, double2* x, PairTab vv, double2* part,
                  int nv, int no, hipStream_t st);
// Z-vector: dg = eps_a - eps_i - Re R[ai,ai]; part = chunked R x - B conj(x)
// (zvec_chunks(n) chunks of n); xn = (rhs + sum part - R[r,r] x) / dg; y = x / dg
void k_zvec_dg(const double2* R, const double* eps, double* dg, int nv, int no, hipStream_t st);
int zvec_chunks(int64_t n);
void k_zvec_gemv(const double2* R, const double2* B, const double2* x, double2* part, int64_t n,
                 hipStream_t st);
void k_zvec_step(const double2* part, const double2* rhs, const double2* R, const double2* x,
                 const double* dg, double2* xn, int64_t n, hipStream_t st);
void k_zvec_precd(const double2* x, const double* dg, double2* y, int64_t n, hipStream_t st);
// every block of a symmetry layout in one launch (gather: C1 -> blocks)
struct BlockSet {
    int nb;
    struct Blk {
        const int* rmap;
        const int* cmap;
        int64_t rows, cols, off;
    } b[32];
};
void k_blockset(const double2* src, const BlockSet& bs, double2* dst, int64_t ld, bool gather,
                hipStream_t st);

// ---- split-BF16 operands and GEMM (rcc_split.hip) ----
// Re, Im and Re+Im as three BF16 slices each, fragment-major, rows padded to
// 128 and k to 128, plus one power-of-two exponent per (row, 128-k chunk,
// variant).
struct SplitOperand {
    int rows = 0, k = 0, rows_pad = 0, nkc = 0, nks = 0;
    int64_t k_pad = 0, plane_bytes = 0, ex_bytes = 0;
    void* planes = nullptr;          // plane_bytes of slice fragments
    signed char* ex = nullptr;       // ex_bytes exponents
};
void split_layout(SplitOperand& S, int rows, int k);
// X: logical rows x k complex operand; k_contiguous: element (r, kk) at
// X[kk + r*ld], else X[r + kk*ld]; conj: use conj(X).
int split_operand(const SplitOperand& S, const double2* X, int64_t ld, bool k_contiguous,
                  bool conj, hipStream_t st);
// The same for the ring form of T2 (rows (a, i), k = (c, k), nv o x nv o),
//   X[(a,i),(c,k)] = s sgn T2[ac, ik] - t1(a,k) t1(c,i),
// computed from the packed T2 (nvvt x noot) and t1 (nv x o; null: no product
// term) as it is sliced, without a full-size intermediate.
int split_ring_operand(const SplitOperand& S, const double2* T2, const double2* t1, double s,
                       int nv, int no, int64_t nvvt, hipStream_t st);
// The same for the F1 expansion of a packed P (nvvt x ncol), F1(a, d + nv col)
// = sgn P[ad, col] (nv x nv ncol): S sliced from F1 (rows a, k = d + nv col)
// or, transposed, from F1^T (rows d + nv col, k = a).
int split_f1_operand(const SplitOperand& S, const double2* P, int nv, int64_t nvvt, int64_t ncol,
                     bool transposed, hipStream_t st);
// C(m x n) = alpha * A . B^T + beta * C with A (m rows) and B (n rows) split
// operands over the same k.  ws (ws_elems complex, may be null) holds the
// partial sums of a K-split.
int zgemm_split(const SplitOperand& A, const SplitOperand& B, double2* C, int64_t ldc,
                double2 alpha, double2 beta, double2* ws, int64_t ws_elems, hipStream_t st);
// C = alpha sum_s W[s] + beta C over nsplit m x n partials W (ld m, stride m n),
// summed as zgemm's K-split reduction
void launch_split_reduce(int m, int n, int nsplit, const double2* W, double2* C, int64_t ldc,
                         double2 alpha, double2 beta, hipStream_t st);

// ---- split-INT8 operands and GEMM (rcc_split_i8.hip) ----
// Re, Im and Re+Im as `digits` signed 7-bit digits per value (fixed point per
// row and 2048-k chunk), fragment-major; exact int32 accumulation.
struct I8Operand {
    int rows = 0, k = 0, digits = 5, rows_pad = 0, nks = 0, nch = 0;
    int64_t plane_bytes = 0, mx_elems = 0;
    void* planes = nullptr;           // plane_bytes
    unsigned long long* mx = nullptr; // mx_elems chunk maxima (bit patterns of doubles)
};
void i8_layout(I8Operand& S, int rows, int k, int digits);
int i8_split(const I8Operand& S, const double2* X, int64_t ld, bool k_contiguous, bool conj,
             hipStream_t st);
int zgemm_i8(const I8Operand& A, const I8Operand& B, double2* C, int64_t ldc, double2 alpha,
             double2 beta, hipStream_t st);
// shapes whose tile height the INT8 autotuner has picked so far
int64_t i8_tuned_shapes();
// the 128 x 128-tile form of zgemm_i8 (rcc_split_i8w.hip), four digits
int launch_i8_wide(int S, int gm, const I8Operand& A, const I8Operand& B, double2* C, int64_t ldc,
                   double2 alpha, double2 beta, hipStream_t st);

// ---- K-split reduction (shared by zgemm_reduce_kernel and the batched
// split reduction, so the two give the same bits) ----
// Lanes per output: lane q of an output's group adds the splits q, q + ways,
// .. in order, then the group combines by a fixed shuffle butterfly.
__host__ __device__ inline int split_ways(int nsplit) {
    return nsplit >= 128 ? 16 : nsplit >= 32 ? 8 : nsplit >= 8 ? 4 : 1;
}

// sum over the nsplit partials w[s * stride] of one output; every lane of
// the group calls it, lane q == 0 gets the sum
__device__ inline double2 split_sum(const double2* __restrict__ w, int nsplit, int64_t stride,
                                    int q, int ways) {
    double sr = 0.0, si = 0.0;
    for (int s = q; s < nsplit; s += ways) {
        const double2 v = w[(int64_t)s * stride];
        sr += v.x;
        si += v.y;
    }
    for (int off = ways >> 1; off > 0; off >>= 1) {
        sr += __shfl_down(sr, off, ways);
        si += __shfl_down(si, off, ways);
    }
    return make_double2(sr, si);
}

}  // namespace rcc
