// Host side of the MI355X relccsd CCSD amplitude solver and its C ABI.
//
// One handle holds, in HBM, the six integral classes in relccsd's packed
// layouts, the constant re-orderings of them that the contractions read
// (built once, since the integrals do not change during the iteration), the
// amplitudes, the intermediates and the DIIS store.  Each per-term function
// mirrors one relccsd routine: the same intermediate, the same packed
// layout, and the same sums (triangular C>D / K>L sums are kept wherever the
// reference keeps them, so the flop count of the ladder and of tau.A is the
// reference's), expressed as FP64-MFMA complex GEMMs plus HBM-bound
// expansion / antisymmetrising kernels.
//
// Device layouts (C1; o occupied, v virtual, nvo = o v, NOOT / NVVT packed
// pair counts):
//   T = [T1 (v x o) | T2 (NVVT x NOOT)], S likewise, TAU (NVVT x NOOT)
//   VVOO_E2 [(CD,L), K] = <cd||kl>              (HOINTM operand)
//   VVOO_F1 [C, (D,KL)] = <cd||kl>              (HVINTM operand)
//   VVOO_R  [(C,K),(D,L)] = <cd||kl>            (HOVINTM, HINTM operand)
//   VOVO_R  [(A,I),(C,K)] = <ak||ci>            (T1EQNS operand)
//   VOVV_Q  [(A,K), C + v D] = <ak||cd>         (HINTM / GVINTM operand)
//   M1 [(C,KL), I] = conj VOOO(CI,KL)           (T1EQN1 operand)
//   MGO [(C,L),(K,I)] = conj <ci||kl>           (GOINTM operand)
//   MH  [L,(K,(C,I))] = conj <ci||lk>           (HINTM operand)
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include "rcc_internal.h"
#include "rcc_blocked.h"
#include "../../include/relccsd_hip.h"

using namespace rcc;

namespace {

constexpr int RED_BLOCKS = 1024;
constexpr int MAXDIM_CAP = 16;
constexpr int RED_SLOTS = 40;
constexpr int RED_E = 32;          // first energy slot in red_dst

struct Buf {
    double2* p = nullptr;
    int64_t n = 0;
};

}  // namespace

// Symmetry-blocked layouts of SYMTAB (ccdriv.F:746-1425).  The solver runs
// in the C1 layout of the whole spinor set (occupied irrep-major, then
// virtual irrep-major -- the order SYMTAB gives the orbital energies,
// ccdriv.F:1392-1402); blocks are scattered into / gathered out of it
// through per-block maps of C1 pair indices.  Off-symmetry elements stay
// exactly zero through the iterations (every product they would receive has
// a zero factor).
enum { SK_OO = 0, SK_VV = 1, SK_VO = 2, SK_OCC = 3, SK_VIR = 4 };

struct SymInfo {
    int nrep = 1;
    std::vector<int> no, nv, io, iv;
    std::vector<int> ptab;                  // pair irrep of (i, j): ptab[i + nrep j], 0-based
    std::vector<std::vector<int>> map[5];   // per block: C1 index of each packed row/column
    std::vector<int64_t> doff[5];           // offset of map[k][b] in dmap
    int* dmap = nullptr;
};

// One logical split operand, in the format of the current precision mode.
struct SOp {
    SplitOperand bf;
    I8Operand i8;
};

struct rcc_handle {
    int device = 0;        // the device current at rcc_create (its stream's device)
    int no = 0, nv = 0;
    int64_t noot = 0, nvvt = 0, nvo = 0, nt1 = 0, nt2 = 0;
    int maxdim = 8, maxit = 30, ntol = 12, precision = 0;
    int gemm_gauss = 1;    // complex products: 1 = three real MFMA products, 0 = four
    hipStream_t st = nullptr;
    std::string err;
    std::vector<void*> allocs;

    int2* oo_tab = nullptr;
    int2* vv_tab = nullptr;
    PairTab oo{}, vv{};
    double* eps = nullptr;
    int* frozen = nullptr;
    bool any_frozen = false;

    double2 *FOO = nullptr, *FVO = nullptr, *FVV = nullptr;
    double2* cls[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int64_t cls_rows[6] = {0}, cls_cols[6] = {0};
    bool derived_ok = false;

    double2 *VVOO_E2 = nullptr, *VVOO_F1 = nullptr, *VVOO_R = nullptr, *VOVO_R = nullptr,
            *VOVV_Q = nullptr, *M1 = nullptr, *MGO = nullptr, *MH = nullptr;

    double2 *T = nullptr, *S = nullptr, *TAU = nullptr;
    double2 *HO = nullptr, *HV = nullptr, *HOV = nullptr, *GO = nullptr, *GV = nullptr;
    double2 *AZ = nullptr, *H = nullptr, *X = nullptr;
    double2 *E2 = nullptr, *F1 = nullptr, *RING = nullptr, *RING2 = nullptr, *HAKCI = nullptr;
    double2 *YP = nullptr, *ZX = nullptr, *WC = nullptr, *Y1 = nullptr, *YAZ = nullptr;
    double2 *SMALL = nullptr;    // o*o + v*o scratch
    double2* ws = nullptr;
    int64_t ws_elems = 0;
    double2* red_part = nullptr;
    double2* red_dst = nullptr;   // RED_SLOTS results: [0,2) norm, [2,2+16) DIIS dots, 32.. energy
    double2* hred = nullptr;      // pinned host copy of red_dst

    std::vector<double2*> diis_err, diis_amp;
    int it = 0, itc = 0;
    std::vector<double> bb2;  // (MAXDIM+1) x (MAXDIM+2), column-major like BB2
    int nit = 0;
    bool tau_ok = false, x_ok = false, ring_ok = false;
    // split operands (precision 1: BF16 slices, 2: INT8 digits): the integral
    // operands are built once per set of integrals, the rest per iteration
    SOp sVVVV, sVOVV, sVVOOR, sTAU, sP, sH, sRING;
    // split-BF16 only: the AINTM, HVINTM and GV . T2 products
    SOp sVVOOA, sVVOOF1, sF1A, sF1B, sGV, sTAUA, sAZ;
    bool split_const_ok = false, split_tau_ok = false, split_ring_ok = false;
    int digits = 5;
    double last_iter_ms = 0.0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // kernel timers: 0 ladder, 1 HINTM ring, 2 T2EQNS ring, 3 BINTM X
    hipEvent_t kev[4][2] = {};
    double kms[4] = {0, 0, 0, 0};
    bool ktimed[4] = {false, false, false, false};

    SymInfo sym;
    double2* stage = nullptr;   // device staging for host blocks
    int64_t stage_n = 0;
    // grouped contractions over irrep blocks (nrep > 1): VVVV and VOVV are
    // also kept in their symmetry-blocked layout; the ladder, VOVV.tau and
    // (for self-inverse irreps) the two rings run as one grouped GEMM each on
    // block copies of their operands
    bool ring_grouped = false;
    double2 *VVVV_s = nullptr, *VOVV_s = nullptr, *TAU_s = nullptr, *S2_s = nullptr,
            *X_s = nullptr, *H_s = nullptr, *RING_s = nullptr, *RING2_s = nullptr,
            *VVOOR_s = nullptr;
    ZGroup *g_ladder = nullptr, *g_x = nullptr, *g_ring1 = nullptr, *g_ring2 = nullptr;
    // host copies of the descriptors (0 ladder, 1 X, 2 HINTM ring, 3 T2EQNS
    // ring), concatenated across handles by rcc_iterate_batch_grouped
    std::vector<ZGroup> hg[4];
    int ng_ladder = 0, ng_x = 0, ng_ring1 = 0, ng_ring2 = 0;
    int64_t nb_ladder = 0, nb_x = 0, nb_ring1 = 0, nb_ring2 = 0;
    bool vvoor_s_ok = false, tau_s_ok = false, ring_s_ok = false;
    // hipGraph of the term sequence HOINTM .. DENOM (replayed each iteration)
    bool use_graph = false, graph_warm = false, graph_ok = false, capturing = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;

    // Lambda equations (C1, FP64): LAM the current Lambda, LAMS the next one
    // (and the output of the adjoint sweep), b* the adjoints of the
    // iteration's intermediates, LW / LB2 the T1EQNS products W = HOV -
    // 2 conj(FVO) and W^T T1 at the converged T
    double2 *LAM = nullptr, *LAMS = nullptr;
    double2 *bTAU = nullptr, *bX = nullptr, *bWC = nullptr, *bRING = nullptr, *bRING2 = nullptr,
            *bH = nullptr, *bE2 = nullptr, *bF1 = nullptr, *bP = nullptr, *ones = nullptr;
    double2 *bHO = nullptr, *bHV = nullptr, *bHOV = nullptr, *bGO = nullptr, *bGV = nullptr,
            *bAZ = nullptr, *bB2 = nullptr, *bW = nullptr, *bFVOc = nullptr, *LW = nullptr,
            *LB2 = nullptr, *bPT2 = nullptr;
    bool lam_alloc = false, lam_set = false;
    double last_lambda_ms = 0.0;
    // symmetric handles: the ladder adjoint TAUb(c) = VVVV(c)^T S2b(c) as one
    // grouped launch over the compound irreps, on block copies
    double2 *S2b_s = nullptr, *bTAU_s = nullptr;
    ZGroup* g_ladder_adj = nullptr;
    int ng_ladder_adj = 0;
    int64_t nb_ladder_adj = 0;
    // non-iterative triples: work buffers and the last call's timing
    T3State* t3 = nullptr;
    T3Timing t3_tm;
    double last_sort_ms = 0.0;     // last integral sort (rcc_read_molfdir / rcc_sort_ints_device)
    // MP2 relaxed density (rcc_mp2_density): the MP2 amplitudes, the density
    // blocks DOO (o x o), DVV (v x v), the X intermediate as CCDENSZ stores it
    // (-X), DVO and DOV (VO vectors), the RPA preconditioner and the Z-vector
    // work (DIIS history, partial sums)
    double2* mp2t = nullptr;
    double2* mp2d = nullptr;
    double* zdg = nullptr;
    double2* zwork = nullptr;
    int64_t zwork_n = 0;
    bool mp2d_ok = false;
    int mp2_nit = 0;
    double mp2_rms = 0.0, last_mp2d_ms = 0.0;
    // phase timers: 0 start, 1 densities done, 2 X done, 3 end; 4/5 around
    // each Z-vector matrix-vector launch, 6/7 around the VOVV x lambda GEMM.
    // mp2_ms: total, densities, X, Z-vector loop, sum of the matrix-vector
    // launches, the VOVV x lambda GEMM
    hipEvent_t mev[8] = {};
    double mp2_ms[6] = {0, 0, 0, 0, 0, 0};

    // symmetry-blocked engine (symmetric handles whose irreps are their own
    // inverse): the iteration's state and work live in its blocks; the C1
    // arrays above serve the paths that run in the C1 layout (Lambda, (T),
    // MP2 density), which take T from the engine first (c1t_stale)
    BlkEngine* bk = nullptr;
    bool bk_const_ok = false;
    bool c1t_stale = false;
    // the C1 working set (classes, their re-orderings, amplitudes and
    // intermediates in the C1 layout) exists; blocked handles allocate it only
    // for a path that runs in the C1 layout
    bool c1_ok = false;
    int64_t diis_len = 0;      // elements per DIIS vector allocated

    double2* T1() { return T; }
    double2* T2() { return T + nt1; }
    double2* S1() { return S; }
    double2* S2() { return S + nt1; }
};

namespace {

#define HIPCHK(h, call)                                                               \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            (h)->err = std::string(#call) + ": " + hipGetErrorString(e_);             \
            return -1;                                                                \
        }                                                                             \
    } while (0)

// symmetry-layout helpers (defined with the layout code further down)
void c1_to_sym_dev(rcc_t* h, int lay, const double2* src, double2* dst);
void sym_to_c1_dev(rcc_t* h, int lay, const double2* src, double2* dst);
int64_t lay_total(const rcc_t* h, int lay);
int run_groups(rcc_t* h, int opb, int tag, const ZGroup* dG, int ng, int64_t nblocks);

int dalloc(rcc_t* h, double2** p, int64_t n) {
    *p = nullptr;
    if (n <= 0) n = 1;
    hipError_t e = hipMalloc((void**)p, (size_t)n * sizeof(double2));
    if (e != hipSuccess) {
        h->err = "hipMalloc of " + std::to_string((double)n * 16.0 / 1e9) + " GB failed: " +
                 hipGetErrorString(e);
        (void)hipGetLastError();     // reported here, not again by a later launch check
        *p = nullptr;
        return -1;
    }
    h->allocs.push_back(*p);
    return 0;
}

template <class T>
int dalloc_raw(rcc_t* h, T** p, int64_t bytes) {
    hipError_t e = hipMalloc((void**)p, (size_t)std::max<int64_t>(bytes, 16));
    if (e != hipSuccess) {
        h->err = std::string("hipMalloc failed: ") + hipGetErrorString(e);
        (void)hipGetLastError();
        return -1;
    }
    h->allocs.push_back(*p);
    return 0;
}

inline double2 cplx(double r, double i) { return make_double2(r, i); }

int gemm(rcc_t* h, int opa, int opb, int m, int n, int k, double2 alpha, const double2* A,
         int64_t lda, const double2* B, int64_t ldb, double2 beta, double2* C, int64_t ldc,
         int batch = 1, int64_t sA = 0, int64_t sB = 0, int64_t sC = 0, int tag = 0) {
    GemmArgs g;
    g.tag = tag;
    g.gauss = h->gemm_gauss;
    g.opA = opa; g.opB = opb; g.m = m; g.n = n; g.k = k;
    g.A = A; g.lda = lda; g.sA = sA;
    g.B = B; g.ldb = ldb; g.sB = sB;
    g.C = C; g.ldc = ldc; g.sC = sC;
    g.alpha = alpha; g.beta = beta; g.batch = batch;
    int rc = zgemm(g, h->ws, h->ws_elems, h->st);
    if (rc != 0) {
        h->err = "zgemm launch failed (" + std::to_string(rc) + ")";
        return -1;
    }
    return 0;
}

void ktick(rcc_t* h, int which, int phase) {
    if (h->capturing) return;   // graph replays are timed as a whole
    (void)hipEventRecord(h->kev[which][phase], h->st);
    if (phase == 1) h->ktimed[which] = true;
}

#define CK(x)                 \
    do {                      \
        if ((x) != 0) return -1; \
    } while (0)

inline int64_t pair_index(int x, int y, int n) { return (int64_t)y * (2 * n - y - 1) / 2 + (x - y - 1); }

// The amplitude vectors [T1 | T2] / [S1 | S2] the iteration works on: the
// blocked engine's (symmetry-blocked layout) or the C1 arrays.
inline int64_t amp_n1(const rcc_t* h) { return h->bk ? blk_size(h->bk, blk::A_T1) : h->nt1; }
inline int64_t amp_n2(const rcc_t* h) { return h->bk ? blk_size(h->bk, blk::A_T2) : h->nt2; }
inline int64_t amp_n(const rcc_t* h) { return amp_n1(h) + amp_n2(h); }
inline double2* amp_T(rcc_t* h) { return h->bk ? blk_ptr(h->bk, blk::A_T) : h->T; }
inline double2* amp_S(rcc_t* h) { return h->bk ? blk_ptr(h->bk, blk::A_S) : h->S; }
inline double2* amp_TAU(rcc_t* h) { return h->bk ? blk_ptr(h->bk, blk::A_TAU) : h->TAU; }
inline const double2* amp_FVO(rcc_t* h) { return h->bk ? blk_ptr(h->bk, blk::A_FVO) : h->FVO; }
inline const double2* amp_VVOO(rcc_t* h) {
    return h->bk ? blk_ptr(h->bk, blk::A_VVOO) : h->cls[RCC_VVOO];
}

// one program of the blocked engine on the handle's stream
int brun(rcc_t* h, int prog) {
    return blk_run(h->bk, prog, h->gemm_gauss, h->st, h->err) == 0 ? 0 : -1;
}

// Builds the derived constant layouts once all classes are present.
int build_derived(rcc_t* h) {
    if (h->derived_ok) return 0;
    for (int c = 0; c < 6; ++c)
        if (!h->cls[c]) {
            h->err = "integral class " + std::to_string(c) + " not set";
            return -1;
        }
    const int o = h->no, v = h->nv;
    const int64_t nvo = h->nvo, noot = h->noot, nvvt = h->nvvt;
    hipStream_t st = h->st;
    k_expand_e2(h->cls[RCC_VVOO], h->VVOO_E2, nvvt, h->oo, o, st);
    k_expand_f1(h->cls[RCC_VVOO], h->VVOO_F1, noot, h->vv, v, st);
    k_expand_ring(h->cls[RCC_VVOO], nullptr, 1.0, h->VVOO_R, h->vv, h->oo, v, o, st);
    if (h->ring_grouped) c1_to_sym_dev(h, RCC_VOVO, h->VVOO_R, h->VVOOR_s);
    k_perm_akci(h->cls[RCC_VOVO], h->VOVO_R, v, o, false, st);
    k_expand_cols(h->cls[RCC_VOVV], h->VOVV_Q, nvo, h->vv, v, st);
    // small VOOO re-orderings through host-built gather maps
    auto gather = [&](double2* out, const std::vector<int64_t>& idx,
                      const std::vector<signed char>& sg) -> int {
        const int64_t n = (int64_t)idx.size();
        if (n == 0) return 0;
        int64_t* didx = nullptr;
        signed char* dsg = nullptr;
        HIPCHK(h, hipMalloc((void**)&didx, n * sizeof(int64_t)));
        HIPCHK(h, hipMalloc((void**)&dsg, n));
        HIPCHK(h, hipMemcpyAsync(didx, idx.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, st));
        HIPCHK(h, hipMemcpyAsync(dsg, sg.data(), n, hipMemcpyHostToDevice, st));
        k_gather(h->cls[RCC_VOOO], didx, dsg, true, out, n, st);
        HIPCHK(h, hipStreamSynchronize(st));
        HIPCHK(h, hipFree(didx));
        HIPCHK(h, hipFree(dsg));
        return 0;
    };
    {   // M1[(c + v*kl) + v*noot*i] = conj VOOO[(c + v*i) + nvo*kl]
        std::vector<int64_t> idx((size_t)v * noot * o);
        std::vector<signed char> sg(idx.size(), 1);
        for (int i = 0; i < o; ++i)
            for (int64_t kl = 0; kl < noot; ++kl)
                for (int c = 0; c < v; ++c)
                    idx[c + v * kl + (int64_t)v * noot * i] = (c + (int64_t)v * i) + nvo * kl;
        CK(gather(h->M1, idx, sg));
    }
    {   // MGO[(c + v*l) + nvo*(k + o*i)] = conj VOOO_E[(c + v*i), (k,l)]
        std::vector<int64_t> idx((size_t)nvo * o * o);
        std::vector<signed char> sg(idx.size(), 1);
        for (int i = 0; i < o; ++i)
            for (int k = 0; k < o; ++k)
                for (int l = 0; l < o; ++l)
                    for (int c = 0; c < v; ++c) {
                        const int64_t e = (c + (int64_t)v * l) + nvo * (k + (int64_t)o * i);
                        if (k == l) { idx[e] = -1; continue; }
                        const int64_t p = k > l ? pair_index(k, l, o) : pair_index(l, k, o);
                        idx[e] = (c + (int64_t)v * i) + nvo * p;
                        sg[e] = k > l ? 1 : -1;
                    }
        CK(gather(h->MGO, idx, sg));
    }
    {   // MH[l + o*(k + o*ci)] = conj VOOO_E[ci, (l,k)]
        std::vector<int64_t> idx((size_t)o * o * nvo);
        std::vector<signed char> sg(idx.size(), 1);
        for (int64_t ci = 0; ci < nvo; ++ci)
            for (int k = 0; k < o; ++k)
                for (int l = 0; l < o; ++l) {
                    const int64_t e = l + (int64_t)o * (k + (int64_t)o * ci);
                    if (k == l) { idx[e] = -1; continue; }
                    const int64_t p = l > k ? pair_index(l, k, o) : pair_index(k, l, o);
                    idx[e] = ci + nvo * p;
                    sg[e] = l > k ? 1 : -1;
                }
        CK(gather(h->MH, idx, sg));
    }
    HIPCHK(h, hipStreamSynchronize(st));
    h->derived_ok = true;
    return 0;
}

int ensure_tau(rcc_t* h) {
    if (!h->tau_ok) {
        k_gettau(h->T1(), h->T2(), h->TAU, 1.0, h->vv, h->oo, h->nv, h->st);
        h->tau_ok = true;
        h->tau_s_ok = false;
    }
    if (h->sym.nrep > 1 && !h->tau_s_ok) {
        c1_to_sym_dev(h, RCC_LAY_T2, h->TAU, h->TAU_s);
        h->tau_s_ok = true;
    }
    return 0;
}

int ensure_ring(rcc_t* h) {
    if (!h->ring_ok) {
        k_expand_ring(h->T2(), nullptr, 1.0, h->RING, h->vv, h->oo, h->nv, h->no, h->st);
        h->ring_ok = true;
        h->ring_s_ok = false;
    }
    return 0;
}

// BINTM's X(AK,IJ) = sum_{C>D} VOVV(AK,CD) TAU(CD,IJ)  (ccintm.F:772-788); also the
// sum that T1EQN2 (cceqns.F:2601-2714) takes the (I,K) diagonal of.
void invalidate_t(rcc_t* h) {
    if (h->bk) blk_invalidate(h->bk);
    h->tau_s_ok = false;
    h->ring_s_ok = false;
    h->tau_ok = false;
    h->x_ok = false;
    h->ring_ok = false;
    h->split_tau_ok = false;
    h->split_ring_ok = false;
}

void dfree(rcc_t* h, void* p) {
    if (!p) return;
    auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
    if (it != h->allocs.end()) h->allocs.erase(it);
    (void)hipFree(p);
}

void free_split_op(rcc_t* h, SOp& S) {
    dfree(h, S.bf.planes);
    dfree(h, S.bf.ex);
    S.bf = SplitOperand();
    dfree(h, S.i8.planes);
    dfree(h, S.i8.mx);
    S.i8 = I8Operand();
}

void free_split(rcc_t* h) {
    for (SOp* S : {&h->sVVVV, &h->sVOVV, &h->sVVOOR, &h->sTAU, &h->sP, &h->sH, &h->sRING,
                   &h->sVVOOA, &h->sVVOOF1, &h->sF1A, &h->sF1B, &h->sGV, &h->sTAUA, &h->sAZ})
        free_split_op(h, *S);
    h->split_const_ok = h->split_tau_ok = h->split_ring_ok = false;
}

// Split X (rows x k; k_contiguous: element (r, kk) at X[kk + r ld], else
// X[r + kk ld]; conj: of conj(X)) into S in the current precision's format.
// The BF16 planes of S for a rows x k operand (allocated on first use).
int bf_alloc(rcc_t* h, SOp& S, int rows, int64_t k) {
    if (!S.bf.planes) {
        split_layout(S.bf, rows, (int)k);
        CK(dalloc_raw(h, (char**)&S.bf.planes, std::max<int64_t>(1, S.bf.plane_bytes)));
        CK(dalloc_raw(h, &S.bf.ex, std::max<int64_t>(1, S.bf.ex_bytes)));
    }
    return 0;
}

int do_split(rcc_t* h, SOp& S, int rows, int64_t k, const double2* X, int64_t ld, bool kcont,
             bool conj) {
    if (h->precision == 1) {
        CK(bf_alloc(h, S, rows, k));
        if (split_operand(S.bf, X, ld, kcont, conj, h->st) != 0) {
            h->err = "split_operand launch failed";
            return -1;
        }
        return 0;
    }
    if (!S.i8.planes || S.i8.digits != h->digits) {
        free_split_op(h, S);
        i8_layout(S.i8, rows, (int)k, h->digits);
        CK(dalloc_raw(h, (char**)&S.i8.planes, S.i8.plane_bytes));
        CK(dalloc_raw(h, &S.i8.mx, S.i8.mx_elems * 8));
    }
    if (i8_split(S.i8, X, ld, kcont, conj, h->st) != 0) {
        h->err = "i8_split launch failed";
        return -1;
    }
    return 0;
}

int sgemm(rcc_t* h, const SOp& A, const SOp& B, double2* C, int64_t ldc, double2 alpha,
          double2 beta) {
    const int rc = h->precision == 1
                       ? zgemm_split(A.bf, B.bf, C, ldc, alpha, beta, h->ws, h->ws_elems, h->st)
                                     : zgemm_i8(A.i8, B.i8, C, ldc, alpha, beta, h->st);
    if (rc != 0) {
        h->err = "split GEMM launch failed";
        return -1;
    }
    return 0;
}

// In split-BF16 mode the compute-bound o^2 v^3 and o^4 v^2 products besides X
// (AINTM's VVOO^H tau, HVINTM's F1 VVOO_F1^H, T2EQNS's GV F1 and tau AZ) run
// as split products too; the INT8 kernels have no K-split for their few
// output tiles.
bool bf_small_products(const rcc_t* h) { return h->precision == 1 && h->sym.nrep == 1; }

// Constant split operands (integrals), built once per set of integrals.
int ensure_split_const(rcc_t* h) {
    if (h->split_const_ok) return 0;
    const int64_t nvo = h->nvo, nvvt = h->nvvt;
    CK(do_split(h, h->sVVVV, (int)nvvt, nvvt, h->cls[RCC_VVVV], nvvt, false, false));
    CK(do_split(h, h->sVOVV, (int)nvo, nvvt, h->cls[RCC_VOVV], nvo, false, false));
    // HINTM ring B operand: B(k=(d,l), n=(c,k)) = conj VVOO_R[(c,k),(d,l)]
    CK(do_split(h, h->sVVOOR, (int)nvo, nvo, h->VVOO_R, nvo, false, true));
    if (bf_small_products(h)) {
        // AINTM A operand: A(ij, k=ab) = conj VVOO[ab, ij]
        CK(do_split(h, h->sVVOOA, (int)h->noot, nvvt, h->cls[RCC_VVOO], nvvt, true, true));
        // HVINTM B operand: B(c, k) = conj VVOO_F1[c, k], k = (b, kl)
        CK(do_split(h, h->sVVOOF1, h->nv, (int64_t)h->nv * h->noot, h->VVOO_F1, h->nv, false,
                    true));
    }
    h->split_const_ok = true;
    return 0;
}

int ensure_split_tau(rcc_t* h) {
    if (h->split_tau_ok) return 0;
    CK(do_split(h, h->sTAU, (int)h->noot, h->nvvt, h->TAU, h->nvvt, true, false));
    h->split_tau_ok = true;
    return 0;
}

// The four dominant products (X = VOVV.tau, the HINTM and T2EQNS rings and
// the ladder) are each split into the work before the product, the product
// and the work after it, so that a batch of handles can run every handle's
// "before" parts, one grouped launch of the product over all handles, and
// the "after" parts (batch_iterate_grouped).
int x_pre(rcc_t* h) {
    CK(ensure_tau(h));
    if (h->precision != 0) {
        CK(ensure_split_const(h));
        CK(ensure_split_tau(h));
    }
    if (h->sym.nrep > 1) k_fill_zero(h->X_s, lay_total(h, RCC_VOOO), h->st);
    return 0;
}

int x_gemm(rcc_t* h) {
    ktick(h, 3, 0);
    if (h->sym.nrep > 1)    // X blocks: VOVV(c) TAU(c), one grouped launch over the irreps
        CK(run_groups(h, OP_N, 0, h->g_x, h->ng_x, h->nb_x));
    else if (h->precision != 0)
        CK(sgemm(h, h->sVOVV, h->sTAU, h->X, h->nvo, cplx(1, 0), cplx(0, 0)));
    else
        CK(gemm(h, OP_N, OP_N, (int)h->nvo, (int)h->noot, (int)h->nvvt, cplx(1, 0),
                h->cls[RCC_VOVV], h->nvo, h->TAU, h->nvvt, cplx(0, 0), h->X, h->nvo));
    ktick(h, 3, 1);
    return 0;
}

int x_post(rcc_t* h) {
    if (h->sym.nrep > 1) {
        k_fill_zero(h->X, h->nvo * h->noot, h->st);
        sym_to_c1_dev(h, RCC_VOOO, h->X_s, h->X);
    }
    h->x_ok = true;
    return 0;
}

int ensure_x(rcc_t* h) {
    if (!h->x_ok) {
        CK(x_pre(h));
        CK(x_gemm(h));
        CK(x_post(h));
    }
    return 0;
}

int prep(rcc_t* h) {
    if (!h->derived_ok) CK(build_derived(h));
    return 0;
}

// ---------------------------------------------------------------- terms

int do_hointm(rcc_t* h) {
    CK(prep(h));
    CK(ensure_tau(h));
    const int o = h->no;
    k_copy((int64_t)o * o, h->FOO, h->HO, h->st);
    k_zero_diag(h->HO, o, h->st);
    k_expand_e2(h->TAU, h->E2, h->nvvt, h->oo, o, h->st);
    const int64_t K = h->nvvt * o;
    return gemm(h, OP_C, OP_N, o, o, (int)K, cplx(1, 0), h->VVOO_E2, K, h->E2, K, cplx(1, 0),
                h->HO, o);
}

int do_hvintm(rcc_t* h) {
    CK(prep(h));
    CK(ensure_tau(h));
    const int v = h->nv;
    k_copy((int64_t)v * v, h->FVV, h->HV, h->st);
    k_zero_diag(h->HV, v, h->st);
    const int64_t K = (int64_t)v * h->noot;
    if (bf_small_products(h) && K > 0) {
        // A = F1(tau), sliced straight from tau
        CK(ensure_split_const(h));
        CK(bf_alloc(h, h->sF1A, v, K));
        if (split_f1_operand(h->sF1A.bf, h->TAU, v, h->nvvt, h->noot, false, h->st) != 0) {
            h->err = "split_f1_operand launch failed";
            return -1;
        }
        return sgemm(h, h->sF1A, h->sVVOOF1, h->HV, v, cplx(-1, 0), cplx(1, 0));
    }
    k_expand_f1(h->TAU, h->F1, h->noot, h->vv, v, h->st);
    return gemm(h, OP_N, OP_C, v, v, (int)K, cplx(-1, 0), h->F1, v, h->VVOO_F1, v, cplx(1, 0),
                h->HV, v);
}

int do_hovintm(rcc_t* h) {
    CK(prep(h));
    k_conj(h->nvo, h->FVO, h->HOV, h->st);
    return gemm(h, OP_C, OP_N, (int)h->nvo, 1, (int)h->nvo, cplx(1, 0), h->VVOO_R, h->nvo,
                h->T1(), h->nvo, cplx(1, 0), h->HOV, h->nvo);
}

int do_t1eqns(rcc_t* h) {
    CK(prep(h));
    CK(ensure_tau(h));
    CK(ensure_x(h));
    CK(ensure_ring(h));
    const int o = h->no, v = h->nv;
    const int64_t nvo = h->nvo;
    hipStream_t st = h->st;
    double2* S1 = h->S1();
    double2* W = h->SMALL;             // v x o
    double2* B2 = h->SMALL + nvo;      // o x o
    k_copy(nvo, h->FVO, S1, st);
    // T(A,K) [conj HOV - 2 F]^H T(C,I)   (cceqns.F:736-752)
    k_conj(nvo, h->FVO, W, st);
    k_axpby(nvo, cplx(1, 0), h->HOV, cplx(-2, 0), W, st);
    CK(gemm(h, OP_T, OP_N, o, o, v, cplx(1, 0), W, v, h->T1(), v, cplx(0, 0), B2, o));
    CK(gemm(h, OP_N, OP_N, v, o, o, cplx(1, 0), h->T1(), v, B2, o, cplx(1, 0), S1, v));
    // HOV(CK) T2(A,C,I,K)   (cceqns.F:764-771)
    CK(gemm(h, OP_N, OP_N, (int)nvo, 1, (int)nvo, cplx(1, 0), h->RING, nvo, h->HOV, nvo,
            cplx(1, 0), S1, nvo));
    // HV.T - T.HO   (cceqns.F:792,796)
    CK(gemm(h, OP_N, OP_N, v, o, v, cplx(1, 0), h->HV, v, h->T1(), v, cplx(1, 0), S1, v));
    CK(gemm(h, OP_N, OP_N, v, o, o, cplx(-1, 0), h->T1(), v, h->HO, o, cplx(1, 0), S1, v));
    // T1EQN1: V(K,L,C,I) TAU(A,C,K,L)   (cceqns.F:2521-2599)
    k_expand_f1(h->TAU, h->F1, h->noot, h->vv, v, st);
    const int64_t K = (int64_t)v * h->noot;
    CK(gemm(h, OP_N, OP_N, v, o, (int)K, cplx(1, 0), h->F1, v, h->M1, K, cplx(1, 0), S1, v));
    // T1EQN2: V(A,K,C,D) TAU(C,D,I,K)   (cceqns.F:2601-2714)
    k_trace_t1eqn2(h->X, S1, h->oo, v, o, st);
    // - W(A,K,C,I) T(C,K)   (cceqns.F:843-855)
    return gemm(h, OP_N, OP_N, (int)nvo, 1, (int)nvo, cplx(-1, 0), h->VOVO_R, nvo, h->T1(), nvo,
                cplx(1, 0), S1, nvo);
}

int do_gointm(rcc_t* h) {
    CK(prep(h));
    const int o = h->no, v = h->nv;
    k_copy((int64_t)o * o, h->HO, h->GO, h->st);
    CK(gemm(h, OP_C, OP_N, o, o, v, cplx(1, 0), h->FVO, v, h->T1(), v, cplx(1, 0), h->GO, o));
    return gemm(h, OP_T, OP_N, o * o, 1, (int)h->nvo, cplx(-1, 0), h->MGO, h->nvo, h->T1(),
                h->nvo, cplx(1, 0), h->GO, (int64_t)o * o);
}

int do_gvintm(rcc_t* h) {
    CK(prep(h));
    const int o = h->no, v = h->nv;
    k_copy((int64_t)v * v, h->HV, h->GV, h->st);
    CK(gemm(h, OP_N, OP_C, v, v, o, cplx(-1, 0), h->T1(), v, h->FVO, v, cplx(1, 0), h->GV, v));
    // HAKCI[(a,k),(c,i)] = - sum_d <ak||cd> t(d,i); its (k = i) trace is the
    // GV term W(AK,CD) T(D,K) (ccintm.F:208-218); HINTM reuses HAKCI.
    CK(gemm(h, OP_N, OP_N, (int)(h->nvo * v), o, v, cplx(-1, 0), h->VOVV_Q, h->nvo * v, h->T1(), v,
            cplx(0, 0), h->HAKCI, h->nvo * v));
    k_trace_gv(h->HAKCI, h->GV, -1.0, v, o, h->st);
    return 0;
}

int do_aintm(rcc_t* h) {
    CK(prep(h));
    CK(ensure_tau(h));
    const int o = h->no, v = h->nv;
    const int64_t noot = h->noot;
    k_copy(noot * noot, h->cls[RCC_OOOO], h->AZ, h->st);
    if (o > 0 && noot > 0) {
        // YAZ[kl + noot*(j + o*i)] = sum_c conj VOOO[(c,i),kl] t(c,j)
        CK(gemm(h, OP_C, OP_N, (int)noot, o, v, cplx(1, 0), h->cls[RCC_VOOO], h->nvo, h->T1(), v,
                cplx(0, 0), h->YAZ, noot, o, v, 0, noot * o));
        k_pack_pij(h->YAZ, h->AZ, noot, 1.0, h->oo, o, h->st);
    }
    if (bf_small_products(h) && noot > 0) {
        // AZ += VVOO^H tau: A = conj VVOO (noot rows), B = the ladder's tau operand
        CK(ensure_split_const(h));
        CK(ensure_split_tau(h));
        return sgemm(h, h->sVVOOA, h->sTAU, h->AZ, noot, cplx(1, 0), cplx(1, 0));
    }
    return gemm(h, OP_C, OP_N, (int)noot, (int)noot, (int)h->nvvt, cplx(1, 0), h->cls[RCC_VVOO],
                h->nvvt, h->TAU, h->nvvt, cplx(1, 0), h->AZ, noot);
}

int hintm_pre(rcc_t* h, bool hakci_has_vovv) {
    CK(prep(h));
    const int o = h->no, v = h->nv;
    const int64_t nvo = h->nvo;
    hipStream_t st = h->st;
    if (!hakci_has_vovv)
        CK(gemm(h, OP_N, OP_N, (int)(nvo * v), o, v, cplx(-1, 0), h->VOVV_Q, nvo * v, h->T1(), v,
                cplx(0, 0), h->HAKCI, nvo * v));
    // - W(AK,CI)  (ccintm.F:263)
    k_axpby(nvo * nvo, cplx(-1, 0), h->cls[RCC_VOVO], cplx(1, 0), h->HAKCI, st);
    // + T(A,L) W(LK,CI)  (ccintm.F:268-273)
    CK(gemm(h, OP_N, OP_N, v, (int)(o * nvo), o, cplx(1, 0), h->T1(), v, h->MH, o, cplx(1, 0),
            h->HAKCI, v));
    // (AK,CI) -> (AI,CK)  (SRT16, ccintm.F:385)
    k_perm_akci(h->HAKCI, h->H, v, o, false, st);
    // + [1/2 T(AD,IL) - T(A,L) T(D,I)] W(KL,CD)  (ccintm.F:393-450)
    if (h->precision == 1 && !h->ring_grouped) {
        // split-BF16: the ring form is sliced straight from T2 and T1
        CK(ensure_split_const(h));
        CK(bf_alloc(h, h->sP, (int)nvo, nvo));
        if (split_ring_operand(h->sP.bf, h->T2(), h->T1(), 0.5, v, o, h->nvvt, st) != 0) {
            h->err = "split_ring_operand launch failed";
            return -1;
        }
        return 0;
    }
    k_expand_ring(h->T2(), h->T1(), 0.5, h->RING2, h->vv, h->oo, v, o, st);
    if (h->ring_grouped) {
        c1_to_sym_dev(h, RCC_VOVO, h->H, h->H_s);
        c1_to_sym_dev(h, RCC_VOVO, h->RING2, h->RING2_s);
    } else if (h->precision != 0) {
        CK(ensure_split_const(h));
        CK(do_split(h, h->sP, (int)nvo, nvo, h->RING2, nvo, false, false));
    }
    return 0;
}

// + [1/2 T(AD,IL) - T(A,L) T(D,I)] W(KL,CD)  (ccintm.F:393-450)
int hintm_gemm(rcc_t* h) {
    const int64_t nvo = h->nvo;
    ktick(h, 1, 0);
    if (h->ring_grouped)
        CK(run_groups(h, OP_C, 2, h->g_ring1, h->ng_ring1, h->nb_ring1));
    else if (h->precision != 0)
        CK(sgemm(h, h->sP, h->sVVOOR, h->H, nvo, cplx(1, 0), cplx(1, 0)));
    else
        CK(gemm(h, OP_N, OP_C, (int)nvo, (int)nvo, (int)nvo, cplx(1, 0), h->RING2, nvo,
                h->VVOO_R, nvo, cplx(1, 0), h->H, nvo, 1, 0, 0, 0, 2));
    ktick(h, 1, 1);
    return 0;
}

int hintm_post(rcc_t* h) {
    if (h->ring_grouped) sym_to_c1_dev(h, RCC_VOVO, h->H_s, h->H);
    return 0;
}

int do_hintm(rcc_t* h, bool hakci_has_vovv) {
    CK(hintm_pre(h, hakci_has_vovv));
    CK(hintm_gemm(h));
    return hintm_post(h);
}

int t2eqns_pre(rcc_t* h) {
    CK(prep(h));
    CK(ensure_tau(h));
    CK(ensure_ring(h));
    const int o = h->no, v = h->nv;
    const int64_t nvo = h->nvo, noot = h->noot, nvvt = h->nvvt;
    hipStream_t st = h->st;
    double2* S2 = h->S2();
    k_copy(h->nt2, h->cls[RCC_VVOO], S2, st);
    // TAU(AB,KL) A(KL,IJ)  (cceqns.F:1013-1026)
    if (bf_small_products(h) && noot > 0) {
        // A = tau (nvvt rows, k = kl), B(n = ij, k = kl) = AZ[kl + ij noot]
        CK(do_split(h, h->sTAUA, (int)nvvt, noot, h->TAU, nvvt, false, false));
        CK(do_split(h, h->sAZ, (int)noot, noot, h->AZ, noot, true, false));
        CK(sgemm(h, h->sTAUA, h->sAZ, S2, nvvt, cplx(1, 0), cplx(1, 0)));
    } else {
        CK(gemm(h, OP_N, OP_N, (int)nvvt, (int)noot, (int)noot, cplx(1, 0), h->TAU, nvvt, h->AZ,
                noot, cplx(1, 0), S2, nvvt));
    }
    if (o > 1) {
        // P(ij) W(AB,CJ) T1(C,I)  (cceqns.F:1086-1139):  YP[AB,i,j] = sum_c conj VOVV[(c,j),AB] t(c,i)
        CK(gemm(h, OP_C, OP_N, (int)nvvt, o, v, cplx(1, 0), h->cls[RCC_VOVV], nvo, h->T1(), v,
                cplx(0, 0), h->YP, nvvt, o, v, 0, nvvt * o));
        // - P(ij) T2(AB,IK) GO(K,J)  (cceqns.F:1051-1056): E2(T2)[(AB,i),k] = t2(AB,k,i)
        k_expand_e2(h->T2(), h->E2, nvvt, h->oo, o, st);
        CK(gemm(h, OP_N, OP_N, (int)(nvvt * o), o, o, cplx(1, 0), h->E2, nvvt * o, h->GO, o,
                cplx(1, 0), h->YP, nvvt * o));
        k_pack_pij(h->YP, S2, nvvt, 1.0, h->oo, o, st);
    }
    // P(ab)P(ij) T1(A,K) W(BK,IC) T1(C,J) and - P(ab) T1(A,K) W(KB,IJ)  (cceqns.F:1158-1221)
    k_axpby(nvo * noot, cplx(-1, 0), h->cls[RCC_VOOO], cplx(0, 0), h->WC, st);
    if (o > 1) {
        CK(gemm(h, OP_N, OP_N, (int)nvo, o, v, cplx(1, 0), h->cls[RCC_VOVO], nvo, h->T1(), v,
                cplx(0, 0), h->Y1, nvo, o, v * nvo, 0, nvo * o));
        k_pack_pij(h->Y1, h->WC, nvo, -1.0, h->oo, o, st);
    }
    // P(ab) GV(A,C) T2(CB,IJ)  (cceqns.F:1183-1184)
    if (bf_small_products(h) && noot > 0) {
        // ZX = GV F1(T2): A = GV (v rows, k = c), B(n = (b, ij), k = c) = F1[c + n v],
        // sliced straight from T2
        CK(do_split(h, h->sGV, v, v, h->GV, v, false, false));
        CK(bf_alloc(h, h->sF1B, (int)(v * noot), v));
        if (split_f1_operand(h->sF1B.bf, h->T2(), v, nvvt, noot, true, st) != 0) {
            h->err = "split_f1_operand launch failed";
            return -1;
        }
        CK(sgemm(h, h->sGV, h->sF1B, h->ZX, v, cplx(1, 0), cplx(0, 0)));
    } else {
        k_expand_f1(h->T2(), h->F1, noot, h->vv, v, st);
        CK(gemm(h, OP_N, OP_N, v, (int)(v * noot), v, cplx(1, 0), h->GV, v, h->F1, v,
                cplx(0, 0), h->ZX, v));
    }
    CK(gemm(h, OP_N, OP_T, v, v, o, cplx(-1, 0), h->T1(), v, h->WC, v, cplx(1, 0), h->ZX, v,
            (int)noot, 0, nvo, (int64_t)v * v));
    k_pack_pab(h->ZX, S2, noot, 1.0, h->vv, v, st);
    // P(ab)P(ij) H(AI,CK) T2(CK,BJ)  (cceqns.F:1243-1263)
    if (h->ring_grouped) {
        c1_to_sym_dev(h, RCC_VOVO, h->H, h->H_s);
        if (!h->ring_s_ok) {
            c1_to_sym_dev(h, RCC_VOVO, h->RING, h->RING_s);
            h->ring_s_ok = true;
        }
        k_fill_zero(h->RING2_s, lay_total(h, RCC_VOVO), st);
    } else if (h->precision != 0) {
        CK(do_split(h, h->sH, (int)nvo, nvo, h->H, nvo, false, false));
        if (!h->split_ring_ok) {
            CK(do_split(h, h->sRING, (int)nvo, nvo, h->RING, nvo, true, false));
            h->split_ring_ok = true;
        }
    }
    return 0;
}

int t2eqns_gemm(rcc_t* h) {
    const int64_t nvo = h->nvo;
    ktick(h, 2, 0);
    if (h->ring_grouped)
        CK(run_groups(h, OP_N, 2, h->g_ring2, h->ng_ring2, h->nb_ring2));
    else if (h->precision != 0)
        CK(sgemm(h, h->sH, h->sRING, h->RING2, nvo, cplx(1, 0), cplx(0, 0)));
    else
        CK(gemm(h, OP_N, OP_N, (int)nvo, (int)nvo, (int)nvo, cplx(1, 0), h->H, nvo, h->RING, nvo,
                cplx(0, 0), h->RING2, nvo, 1, 0, 0, 0, 2));
    ktick(h, 2, 1);
    return 0;
}

int t2eqns_post(rcc_t* h) {
    if (h->ring_grouped) {
        k_fill_zero(h->RING2, h->nvo * h->nvo, h->st);
        sym_to_c1_dev(h, RCC_VOVO, h->RING2_s, h->RING2);
    }
    k_pack_pabij(h->RING2, h->S2(), h->vv, h->oo, h->nv, h->no, h->st);
    return 0;
}

int do_t2eqns(rcc_t* h) {
    CK(t2eqns_pre(h));
    CK(t2eqns_gemm(h));
    return t2eqns_post(h);
}

int bintm_pre(rcc_t* h) {
    CK(prep(h));
    CK(ensure_tau(h));
    CK(ensure_x(h));
    if (h->sym.nrep > 1) c1_to_sym_dev(h, RCC_LAY_T2, h->S2(), h->S2_s);
    else if (h->precision != 0) {
        CK(ensure_split_const(h));
        CK(ensure_split_tau(h));
    }
    return 0;
}

// ladder: S(AB,IJ) += V(AB,CD) TAU(CD,IJ), C>D  (ccintm.F:700-722); with
// symmetry S2(c) += VVVV(c) TAU(c) over the compound irreps, one grouped launch
int bintm_gemm(rcc_t* h) {
    const int64_t noot = h->noot, nvvt = h->nvvt;
    ktick(h, 0, 0);
    if (h->sym.nrep > 1)
        CK(run_groups(h, OP_N, 1, h->g_ladder, h->ng_ladder, h->nb_ladder));
    else if (h->precision != 0)
        CK(sgemm(h, h->sVVVV, h->sTAU, h->S2(), nvvt, cplx(1, 0), cplx(1, 0)));
    else
        CK(gemm(h, OP_N, OP_N, (int)nvvt, (int)noot, (int)nvvt, cplx(1, 0), h->cls[RCC_VVVV], nvvt,
                h->TAU, nvvt, cplx(1, 0), h->S2(), nvvt, 1, 0, 0, 0, 1));
    ktick(h, 0, 1);
    return 0;
}

int bintm_post(rcc_t* h) {
    const int o = h->no, v = h->nv;
    const int64_t nvo = h->nvo, noot = h->noot;
    double2* S2 = h->S2();
    if (h->sym.nrep > 1) sym_to_c1_dev(h, RCC_LAY_T2, h->S2_s, S2);
    // - P(ab) T(B,K) X(AK,IJ)  (ccintm.F:760-829)
    CK(gemm(h, OP_N, OP_T, v, v, o, cplx(-1, 0), h->X, v, h->T1(), v, cplx(0, 0), h->ZX, v,
            (int)noot, nvo, 0, (int64_t)v * v));
    k_pack_pab(h->ZX, S2, noot, 1.0, h->vv, v, h->st);
    return 0;
}

int do_bintm(rcc_t* h) {
    CK(bintm_pre(h));
    CK(bintm_gemm(h));
    return bintm_post(h);
}

// deterministic complex reductions
int zdot_multi(rcc_t* h, const std::vector<const double2*>& xs, const double2* y, int64_t n,
               std::vector<double2>& out) {
    const int nx = (int)xs.size();
    out.assign(nx, cplx(0, 0));
    for (int off = 0; off < nx; off += MAXDIM_CAP) {
        const int cnt = std::min(MAXDIM_CAP, nx - off);
        k_zdotc_multi(xs.data() + off, cnt, y, n, h->red_dst, h->red_part, RED_BLOCKS, h->st);
        HIPCHK(h, hipMemcpyAsync(out.data() + off, h->red_dst, cnt * sizeof(double2),
                                 hipMemcpyDeviceToHost, h->st));
    }
    HIPCHK(h, hipStreamSynchronize(h->st));
    return 0;
}

// FLIN (ccgenp.F:666-740) restated; a is IDIM x (n+nr) column-major
double dott(const double* a, int na, const double* b, int nb, int n) {
    double d = 0.0;
    for (int i = 0; i < n; ++i) d = d + a[(int64_t)i * na] * b[(int64_t)i * nb];
    return d;
}

double flin(double* A, int idim, int n, int nr) {
    auto a = [&](int i, int j) -> double& { return A[i + (int64_t)j * idim]; };
    const int jmax = n + nr;
    double sign = 1.0, d = 0.0;
    for (int m = 0; m < n; ++m) {
        double temp = 0.0;
        int imax = m;
        for (int i = m; i < n; ++i) {
            if (m > 0) a(i, m) = a(i, m) - dott(&a(i, 0), idim, &a(0, m), 1, m);
            const double aval = a(i, m);
            if (std::fabs(aval) <= temp) continue;
            temp = std::fabs(aval);
            imax = i;
        }
        if (temp <= 0.0) return 0.0;
        if (imax != m) {
            sign = -sign;
            for (int j = 0; j < jmax; ++j) std::swap(a(m, j), a(imax, j));
        }
        const int jj = m + 1;
        if (jj >= jmax) continue;
        if (m == 0) {
            for (int j = jj; j < jmax; ++j) a(0, j) = a(0, j) / a(0, 0);
            d = a(0, 0);
            continue;
        }
        for (int j = jj; j < jmax; ++j) a(m, j) = (a(m, j) - dott(&a(m, 0), idim, &a(0, j), 1, m)) / a(m, m);
        d = d * a(m, m);
    }
    for (int i = 0; i < nr; ++i) {
        const int npi = n + i;
        for (int k = 2; k <= n; ++k) {
            const int j = n - k;
            a(j, npi) = a(j, npi) - dott(&a(j, j + 1), idim, &a(j + 1, npi), 1, k - 1);
        }
    }
    return d * sign;
}

// DIISX (cceqns.F:1985-2178) on device vectors [S1|S2] vs [T1|T2].
// DIISX (cceqns.F:1985-2178) in two halves around one host synchronisation:
// diis_begin enqueues the error vector, its norms and the B-row dots with an
// asynchronous copy of the results to pinned host memory; diis_end (after the
// stream has drained) forms the B matrix, solves it (FLIN) and enqueues the
// extrapolation.
// The circular store position of this DIISX call (cceqns.F:2022-2026).
int diis_advance(rcc_t* h) {
    const int md = h->maxdim;
    h->it = std::min(h->it + 1, md);
    h->itc = h->itc + 1;
    if (h->itc > md) h->itc = 1;
    return h->itc - 1;
}

int diis_begin(rcc_t* h, const double2* cur = nullptr, const double2* nw = nullptr,
               int64_t n = -1) {
    if (!cur) cur = amp_T(h);
    if (!nw) nw = amp_S(h);
    if (n < 0) n = amp_n(h);
    hipStream_t st = h->st;
    const int c = diis_advance(h);
    double2* d = h->diis_err[c];
    // d = S - T ; store S
    HIPCHK(h, hipMemcpyAsync(d, nw, n * 16, hipMemcpyDeviceToDevice, st));
    k_axpby(n, cplx(-1, 0), cur, cplx(1, 0), d, st);
    HIPCHK(h, hipMemcpyAsync(h->diis_amp[c], nw, n * 16, hipMemcpyDeviceToDevice, st));
    // RMS = sqrt(| sum |d|^2 - |sum d|^2 |)
    k_sum_and_norm(d, n, h->red_dst, h->red_part, RED_BLOCKS, st);
    // B row: Re + Im of <d_k | d_ITC>
    std::vector<const double2*> xs;
    for (int k = 0; k < h->it; ++k) xs.push_back(h->diis_err[k]);
    k_zdotc_multi(xs.data(), (int)xs.size(), d, n, h->red_dst + 2, h->red_part, RED_BLOCKS, st);
    HIPCHK(h, hipMemcpyAsync(h->hred, h->red_dst, (2 + xs.size()) * sizeof(double2),
                             hipMemcpyDeviceToHost, st));
    return 0;
}

// DIISX's host part on the reductions in hred: B row, scaling, bordered
// system, FLIN; the extrapolation coefficients.
int diis_coefs(rcc_t* h, double* rms_out, std::vector<double>& coef) {
    const int md = h->maxdim;
    const int c = h->itc - 1;
    const double2* sn = h->hred;
    const double rms = std::sqrt(std::fabs(sn[1].x - (sn[0].x * sn[0].x + sn[0].y * sn[0].y)));
    std::vector<double2> dots(h->hred + 2, h->hred + 2 + h->it);
    const int ld = md + 1;
    auto BB2 = [&](int i, int j) -> double& { return h->bb2[i + (int64_t)j * ld]; };
    for (int k = 0; k < h->it; ++k) {
        const double vv = dots[k].x + dots[k].y;
        BB2(k, c) = vv;
        BB2(c, k) = vv;
    }
    const int it = h->it;
    std::vector<double> bb((size_t)ld * (md + 2), 0.0);
    auto BB = [&](int i, int j) -> double& { return bb[i + (int64_t)j * ld]; };
    for (int i = 0; i < it; ++i)
        for (int j = 0; j < it; ++j) BB(i, j) = BB2(i, j);
    double xm = std::fabs(BB(0, 0));
    for (int i = 0; i < it; ++i)
        for (int j = 0; j <= i; ++j) xm = std::max(xm, std::fabs(BB(i, j)));
    xm = 1.0 / xm;
    for (int i = 0; i < it; ++i)
        for (int j = 0; j < it; ++j) BB(i, j) *= xm;
    for (int i = 0; i < it; ++i) {
        BB(i, it) = -1.0;
        BB(it, i) = -1.0;
    }
    BB(it, it) = 0.0;
    for (int i = 0; i < it; ++i) BB(i, it + 1) = 0.0;
    BB(it, it + 1) = -1.0;
    flin(bb.data(), ld, it + 1, 1);
    coef.assign(it, 0.0);
    for (int i = 0; i < it; ++i) coef[i] = BB(i, it + 1);
    if (rms_out) *rms_out = rms;
    return 0;
}

int diis_end(rcc_t* h, double* rms_out, double2* nw = nullptr, int64_t n = -1) {
    if (!nw) nw = amp_S(h);
    if (n < 0) n = amp_n(h);
    std::vector<double> coef;
    CK(diis_coefs(h, rms_out, coef));
    std::vector<const double2*> amps;
    for (int k = 0; k < h->it; ++k) amps.push_back(h->diis_amp[k]);
    k_lincomb(amps.data(), coef.data(), h->it, nw, n, h->st);
    return 0;
}

int do_diis(rcc_t* h, double* rms_out) {
    CK(diis_begin(h));
    HIPCHK(h, hipStreamSynchronize(h->st));
    return diis_end(h, rms_out);
}

// CCENG (cceqns.F:2180-2270) on S: Re[zdotc(FVO,S1) + zdotc(VVOO,tau(S))]
// CCENG (cceqns.F:2180-2270) on S, split like DIISX around one synchronisation.
int energy_begin(rcc_t* h) {
    if (h->bk) {
        CK(brun(h, blk::P_TAUS));
        const double2* fvo = amp_FVO(h);
        const double2* vvoo = amp_VVOO(h);
        k_zdotc_multi(&fvo, 1, amp_S(h), amp_n1(h), h->red_dst + RED_E, h->red_part, RED_BLOCKS,
                      h->st);
        k_zdotc_multi(&vvoo, 1, amp_TAU(h), amp_n2(h), h->red_dst + RED_E + 1, h->red_part,
                      RED_BLOCKS, h->st);
        HIPCHK(h, hipMemcpyAsync(h->hred + RED_E, h->red_dst + RED_E, 2 * sizeof(double2),
                                 hipMemcpyDeviceToHost, h->st));
        return 0;
    }
    k_gettau(h->S1(), h->S2(), h->TAU, 1.0, h->vv, h->oo, h->nv, h->st);
    h->tau_ok = false;  // TAU now holds tau(S), not tau(T)
    const double2* fvo = h->FVO;
    const double2* vvoo = h->cls[RCC_VVOO];
    k_zdotc_multi(&fvo, 1, h->S1(), h->nt1, h->red_dst + RED_E, h->red_part, RED_BLOCKS, h->st);
    k_zdotc_multi(&vvoo, 1, h->TAU, h->nt2, h->red_dst + RED_E + 1, h->red_part, RED_BLOCKS,
                  h->st);
    HIPCHK(h, hipMemcpyAsync(h->hred + RED_E, h->red_dst + RED_E, 2 * sizeof(double2),
                             hipMemcpyDeviceToHost, h->st));
    return 0;
}

void energy_end(rcc_t* h, double* ecc, double* eimag) {
    const double2 a = h->hred[RED_E], b = h->hred[RED_E + 1];
    if (ecc) *ecc = a.x + b.x;
    if (eimag) *eimag = a.y + b.y;
}

int do_energy(rcc_t* h, double* ecc, double* eimag) {
    CK(energy_begin(h));
    HIPCHK(h, hipStreamSynchronize(h->st));
    energy_end(h, ecc, eimag);
    return 0;
}

int do_t1diag(rcc_t* h, const double2* t1, int64_t n1, double* diag) {
    std::vector<double2> a;
    CK(zdot_multi(h, {t1}, t1, n1, a));
    *diag = std::sqrt(a[0].x / std::max(1, h->no));
    return 0;
}

// Engine constants: the class and Fock blocks, the orbital-energy
// denominators and the integral re-orderings of the program, from the
// handle's arrays (again after any of them changes).
int bprep(rcc_t* h) {
    if (h->bk_const_ok) return 0;
    BlkEngine* E = h->bk;
    c1_to_sym_dev(h, RCC_LAY_FOO, h->FOO, blk_ptr(E, blk::A_FOO));
    c1_to_sym_dev(h, RCC_LAY_T1, h->FVO, blk_ptr(E, blk::A_FVO));
    c1_to_sym_dev(h, RCC_LAY_FVV, h->FVV, blk_ptr(E, blk::A_FVV));
    std::vector<double> eps(h->no + h->nv);
    HIPCHK(h, hipMemcpyAsync(eps.data(), h->eps, eps.size() * 8, hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    if (blk_set_eps(E, eps.data(), h->err) != 0) return -1;
    CK(brun(h, blk::P_CONST));
    HIPCHK(h, hipStreamSynchronize(h->st));
    h->bk_const_ok = true;
    return 0;
}

// The C1 amplitudes from the engine's, for the paths that run in the C1
// layout (Lambda, (T)).
int c1_sync_t(rcc_t* h) {
    if (!h->bk || !h->c1t_stale) return 0;
    k_fill_zero(h->T, h->nt1 + h->nt2, h->st);
    sym_to_c1_dev(h, RCC_LAY_T1, blk_ptr(h->bk, blk::A_T1), h->T1());
    sym_to_c1_dev(h, RCC_LAY_T2, blk_ptr(h->bk, blk::A_T2), h->T2());
    h->tau_ok = h->x_ok = h->ring_ok = false;
    h->tau_s_ok = h->ring_s_ok = false;
    h->c1t_stale = false;
    return 0;
}

// the blocked term sequence, HOINTM .. BINTM, DENOM + ZCORE
int blk_terms(rcc_t* h) {
    CK(bprep(h));
    for (int p : {blk::P_HOINTM, blk::P_HVINTM, blk::P_HOVINTM, blk::P_T1EQNS, blk::P_GOINTM,
                  blk::P_GVINTM, blk::P_AINTM, blk::P_HINTM, blk::P_T2EQNS, blk::P_BINTM,
                  blk::P_DENOM})
        CK(brun(h, p));
    return 0;
}

int run_terms_c1(rcc_t* h);

// HOINTM .. BINTM, DENOM: device work only, no host synchronisation.
int run_terms(rcc_t* h) { return h->bk ? blk_terms(h) : run_terms_c1(h); }

int run_terms_c1(rcc_t* h) {
    CK(do_hointm(h));
    CK(do_hvintm(h));
    CK(do_hovintm(h));
    CK(do_t1eqns(h));
    CK(do_gointm(h));
    CK(do_gvintm(h));
    CK(do_aintm(h));
    CK(do_hintm(h, true));
    CK(do_t2eqns(h));
    CK(do_bintm(h));
    k_denom(h->S1(), h->S2(), h->eps, h->any_frozen ? h->frozen : nullptr, h->vv, h->oo, h->nv,
            h->no, h->st);
    return 0;
}

void drop_graph(rcc_t* h) {
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->graph_exec = nullptr;
    h->graph = nullptr;
    h->graph_ok = false;
    h->graph_warm = false;
}

// The term phase of one iteration, enqueued on the handle's stream.  With the
// graph enabled, the first call runs eagerly (so every buffer the sequence
// needs exists), the second captures the sequence into a hipGraph, and every
// call after that replays it with one launch.  The sequence is the same each
// iteration: it starts from invalidated intermediates and fixed buffers.
int terms_phase(rcc_t* h) {
    CK(h->bk ? bprep(h) : prep(h));
    if (!h->use_graph) return run_terms(h);
    if (h->graph_ok) {
        HIPCHK(h, hipGraphLaunch(h->graph_exec, h->st));
        return 0;
    }
    if (!h->graph_warm) {
        h->graph_warm = true;
        return run_terms(h);
    }
    HIPCHK(h, hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
    h->capturing = true;
    const int rc = run_terms(h);
    h->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->st, &g);
    if (rc != 0 || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rc == 0) h->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
        return -1;
    }
    h->graph = g;
    HIPCHK(h, hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0));
    h->graph_ok = true;
    HIPCHK(h, hipGraphLaunch(h->graph_exec, h->st));
    return 0;
}

// DIISX, CCENG, S -> T and the timing of the iteration (host-synchronous).
// The rest of an iteration in three stages separated by host synchronisation
// of the stream (done by the caller, so a batch synchronises once per stage):
// finish_a enqueues DIISX's reductions; finish_b solves DIISX, enqueues the
// extrapolation, CCENG's reductions and S -> T; finish_c reads the energy.
int finish_a(rcc_t* h) { return diis_begin(h); }

int finish_b(rcc_t* h, double* rms) {
    CK(diis_end(h, rms));
    CK(energy_begin(h));
    HIPCHK(h, hipMemcpyAsync(amp_T(h), amp_S(h), amp_n(h) * 16, hipMemcpyDeviceToDevice, h->st));
    if (h->bk) h->c1t_stale = true;
    HIPCHK(h, hipEventRecord(h->ev1, h->st));
    return 0;
}

int finish_c(rcc_t* h, double* ecc) {
    energy_end(h, ecc, nullptr);
    invalidate_t(h);
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_iter_ms = ms;
    for (int w = 0; w < 4; ++w) {
        float km = 0.f;
        if (h->ktimed[w] && hipEventElapsedTime(&km, h->kev[w][0], h->kev[w][1]) == hipSuccess)
            h->kms[w] = km;
        h->ktimed[w] = false;
    }
    h->nit += 1;
    return 0;
}

int do_iteration(rcc_t* h, double* ecc, double* rms) {
    HIPCHK(h, hipEventRecord(h->ev0, h->st));
    CK(terms_phase(h));
    CK(finish_a(h));
    HIPCHK(h, hipStreamSynchronize(h->st));
    CK(finish_b(h, rms));
    HIPCHK(h, hipStreamSynchronize(h->st));
    return finish_c(h, ecc);
}

int alloc_c1_work(rcc_t* h);

// Arrays every handle needs: Fock blocks, orbital energies, frozen flags,
// the product workspace and the reduction buffers; the C1 working set too
// unless the blocked engine runs the iteration.
int alloc_all(rcc_t* h) {
    const int o = h->no, v = h->nv;
    const int64_t nvo = h->nvo, noot = h->noot, nvvt = h->nvvt;
    CK(dalloc(h, &h->FOO, (int64_t)o * o));
    CK(dalloc(h, &h->FVO, nvo));
    CK(dalloc(h, &h->FVV, (int64_t)v * v));
    CK(dalloc_raw(h, &h->eps, (int64_t)(o + v) * 8));
    CK(dalloc_raw(h, &h->frozen, (int64_t)std::max(o, 1) * 4));
    h->cls_rows[RCC_OOOO] = noot; h->cls_cols[RCC_OOOO] = noot;
    h->cls_rows[RCC_VOOO] = nvo;  h->cls_cols[RCC_VOOO] = noot;
    h->cls_rows[RCC_VVOO] = nvvt; h->cls_cols[RCC_VVOO] = noot;
    h->cls_rows[RCC_VOVO] = nvo;  h->cls_cols[RCC_VOVO] = nvo;
    h->cls_rows[RCC_VOVV] = nvo;  h->cls_cols[RCC_VOVV] = nvvt;
    h->cls_rows[RCC_VVVV] = nvvt; h->cls_cols[RCC_VVVV] = nvvt;
    // K-split partial sums: room for a four-way split of the C1 ladder's output
    // (nvvt x noot), which the tuner may pick to fill the last round of tiles
    // (with symmetry the ladder always runs on the irrep blocks)
    h->ws_elems = std::max<int64_t>({(int64_t)1 << 22,
                                     std::min<int64_t>((int64_t)256 * 1024 * 1024 / 16,
                                                       64 * (int64_t)v * v + (int64_t)256 * o * o),
                                     h->sym.nrep == 1 ? 4 * nvvt * noot : 0});
    CK(dalloc(h, &h->ws, h->ws_elems));
    CK(dalloc(h, &h->red_part, (int64_t)MAXDIM_CAP * RED_BLOCKS));
    CK(dalloc(h, &h->red_dst, RED_SLOTS));
    HIPCHK(h, hipHostMalloc((void**)&h->hred, RED_SLOTS * sizeof(double2), hipHostMallocDefault));
    return h->bk ? 0 : alloc_c1_work(h);
}

// The C1 working set: the classes (VVVV of a symmetric handle only as its
// blocks), their derived layouts, T, S and the intermediates.
int alloc_c1_work(rcc_t* h) {
    if (h->c1_ok) return 0;
    const int o = h->no, v = h->nv;
    const int64_t nvo = h->nvo, noot = h->noot, nvvt = h->nvvt;
    for (int c = 0; c < 6; ++c) {
        // with symmetry the VVVV ladder runs on the blocks only (grouped GEMM)
        const bool blocks_only = c == RCC_VVVV && h->sym.nrep > 1;
        CK(dalloc(h, &h->cls[c], blocks_only ? 1 : h->cls_rows[c] * h->cls_cols[c]));
        // off-symmetry blocks of the C1 arrays are never written
        if (h->sym.nrep > 1 && !blocks_only)
            k_fill_zero(h->cls[c], h->cls_rows[c] * h->cls_cols[c], h->st);
    }
    CK(dalloc(h, &h->VVOO_E2, nvvt * o * o));
    CK(dalloc(h, &h->VVOO_F1, (int64_t)v * v * noot));
    CK(dalloc(h, &h->VVOO_R, nvo * nvo));
    CK(dalloc(h, &h->VOVO_R, nvo * nvo));
    CK(dalloc(h, &h->VOVV_Q, nvo * v * v));
    CK(dalloc(h, &h->M1, (int64_t)v * noot * o));
    CK(dalloc(h, &h->MGO, nvo * o * o));
    CK(dalloc(h, &h->MH, (int64_t)o * o * nvo));
    const int64_t namp = h->nt1 + h->nt2;
    CK(dalloc(h, &h->T, namp));
    CK(dalloc(h, &h->S, namp));
    CK(dalloc(h, &h->TAU, h->nt2));
    CK(dalloc(h, &h->HO, (int64_t)o * o));
    CK(dalloc(h, &h->HV, (int64_t)v * v));
    CK(dalloc(h, &h->HOV, nvo));
    CK(dalloc(h, &h->GO, (int64_t)o * o));
    CK(dalloc(h, &h->GV, (int64_t)v * v));
    CK(dalloc(h, &h->AZ, noot * noot));
    CK(dalloc(h, &h->H, nvo * nvo));
    CK(dalloc(h, &h->X, nvo * noot));
    CK(dalloc(h, &h->E2, nvvt * o * o));
    CK(dalloc(h, &h->F1, (int64_t)v * v * noot));
    CK(dalloc(h, &h->RING, nvo * nvo));
    CK(dalloc(h, &h->RING2, nvo * nvo));
    CK(dalloc(h, &h->HAKCI, nvo * nvo));
    CK(dalloc(h, &h->YP, nvvt * o * o));
    CK(dalloc(h, &h->ZX, (int64_t)v * v * noot));
    CK(dalloc(h, &h->WC, nvo * noot));
    CK(dalloc(h, &h->Y1, nvo * o * o));
    CK(dalloc(h, &h->YAZ, noot * o * o));
    CK(dalloc(h, &h->SMALL, (int64_t)o * o + nvo));
    HIPCHK(h, hipStreamSynchronize(h->st));
    h->c1_ok = true;
    return 0;
}

// DIIS store of MAXDIM error and amplitude vectors of `need` elements (the
// iteration's amplitude vector by default; the Lambda solve asks for C1 ones)
int alloc_diis(rcc_t* h, int64_t need = -1) {
    if (need < 0) need = h->bk ? amp_n(h) : h->nt1 + h->nt2;
    if (need > h->diis_len) {
        for (double2* p : h->diis_err) dfree(h, p);
        for (double2* p : h->diis_amp) dfree(h, p);
        h->diis_err.clear();
        h->diis_amp.clear();
        h->diis_len = need;
    }
    while ((int)h->diis_err.size() < h->maxdim) {
        double2 *e = nullptr, *a = nullptr;
        CK(dalloc(h, &e, h->diis_len));
        CK(dalloc(h, &a, h->diis_len));
        h->diis_err.push_back(e);
        h->diis_amp.push_back(a);
    }
    h->bb2.assign((size_t)(h->maxdim + 1) * (h->maxdim + 2), 0.0);
    h->it = 0;
    h->itc = 0;
    return 0;
}

int upload(rcc_t* h, double2* dst, const void* src, int64_t n, hipMemcpyKind kind) {
    HIPCHK(h, hipMemcpyAsync(dst, src, (size_t)n * 16, kind, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    return 0;
}

// ---- symmetry layouts ----
// Build the pair orderings of SYMTAB: packed pairs of compound irrep c in
// sub-blocks (i >= j, ptab(i, j) = c) ordered by j (IIOOT / IIVVT,
// ccdriv.F:1071-1082), J outer and I inner (I > J on the diagonal; DENOM
// cceqns.F:2408-2448); VO pairs in sub-blocks (a, k) ordered by k (IIVO,
// ccdriv.F:1050-1057), K outer, A inner.
int build_sym(rcc_t* h, const rcc_sym_t* sy) {
    SymInfo& S = h->sym;
    const int n = sy->nrep;
    S.nrep = n;
    S.no.assign(sy->no, sy->no + n);
    S.nv.assign(sy->nv, sy->nv + n);
    S.io.assign(n + 1, 0);
    S.iv.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        S.io[i + 1] = S.io[i] + S.no[i];
        S.iv[i + 1] = S.iv[i] + S.nv[i];
    }
    S.ptab.assign((size_t)n * n, 0);
    if (n > 1)
        for (int i = 0; i < n * n; ++i) S.ptab[i] = sy->multb[i] - 1;
    // partner[j + n c] = the i with ptab(i, j) = c; the table must be a
    // symmetric Latin square (an abelian group's)
    std::vector<int> partner((size_t)n * n, -1);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const int c = S.ptab[i + n * j];
            if (c < 0 || c >= n || partner[j + n * c] >= 0 || S.ptab[j + n * i] != c) {
                h->err = "rcc_create: multb is not the pair-irrep table of an abelian group";
                return -1;
            }
            partner[j + n * c] = i;
        }
    const int64_t NO = S.io[n], NV = S.iv[n];
    auto c1pair = [](int64_t x, int64_t y, int64_t m) {
        return y * (2 * m - y - 1) / 2 + (x - y - 1);
    };
    for (int k = 0; k < 5; ++k) S.map[k].assign(n, {});
    for (int c = 0; c < n; ++c) {
        for (int pass = 0; pass < 2; ++pass) {
            const std::vector<int>& cnt = pass == 0 ? S.no : S.nv;
            const std::vector<int>& off = pass == 0 ? S.io : S.iv;
            const int64_t tot = pass == 0 ? NO : NV;
            std::vector<int>& m = S.map[pass == 0 ? SK_OO : SK_VV][c];
            for (int jr = 0; jr < n; ++jr) {
                const int ir = partner[jr + n * c];
                if (ir < jr) continue;
                for (int j = 0; j < cnt[jr]; ++j)
                    for (int i = ir == jr ? j + 1 : 0; i < cnt[ir]; ++i)
                        m.push_back((int)c1pair(off[ir] + i, off[jr] + j, tot));
            }
        }
        std::vector<int>& vo = S.map[SK_VO][c];
        for (int kr = 0; kr < n; ++kr) {
            const int ar = partner[kr + n * c];
            for (int k = 0; k < S.no[kr]; ++k)
                for (int a = 0; a < S.nv[ar]; ++a)
                    vo.push_back((int)((S.io[kr] + k) * NV + S.iv[ar] + a));
        }
        for (int i = 0; i < S.no[c]; ++i) S.map[SK_OCC][c].push_back(S.io[c] + i);
        for (int a = 0; a < S.nv[c]; ++a) S.map[SK_VIR][c].push_back(S.iv[c] + a);
    }
    int64_t tot = 0;
    for (int k = 0; k < 5; ++k) {
        S.doff[k].assign(n, 0);
        for (int c = 0; c < n; ++c) {
            S.doff[k][c] = tot;
            tot += (int64_t)S.map[k][c].size();
        }
    }
    CK(dalloc_raw(h, &S.dmap, std::max<int64_t>(tot, 1) * 4));
    std::vector<int> all;
    all.reserve(tot);
    for (int k = 0; k < 5; ++k)
        for (int c = 0; c < n; ++c) all.insert(all.end(), S.map[k][c].begin(), S.map[k][c].end());
    if (tot) HIPCHK(h, hipMemcpy(S.dmap, all.data(), tot * 4, hipMemcpyHostToDevice));
    return 0;
}

// Layout ids: the six classes (RCC_OOOO..RCC_VVVV) and RCC_LAY_T1 / T2 /
// FOO / FVV: row and column kinds of each and the C1 leading dimension.
bool lay_spec(const rcc_t* h, int lay, int* rk, int* ck, int64_t* ld) {
    switch (lay) {
        case RCC_OOOO: *rk = SK_OO; *ck = SK_OO; *ld = h->noot; return true;
        case RCC_VOOO: *rk = SK_VO; *ck = SK_OO; *ld = h->nvo; return true;
        case RCC_VVOO: *rk = SK_VV; *ck = SK_OO; *ld = h->nvvt; return true;
        case RCC_VOVO: *rk = SK_VO; *ck = SK_VO; *ld = h->nvo; return true;
        case RCC_VOVV: *rk = SK_VO; *ck = SK_VV; *ld = h->nvo; return true;
        case RCC_VVVV: *rk = SK_VV; *ck = SK_VV; *ld = h->nvvt; return true;
        case RCC_LAY_T1: *rk = SK_VIR; *ck = SK_OCC; *ld = h->nv; return true;
        case RCC_LAY_T2: *rk = SK_VV; *ck = SK_OO; *ld = h->nvvt; return true;
        case RCC_LAY_FOO: *rk = SK_OCC; *ck = SK_OCC; *ld = h->no; return true;
        case RCC_LAY_FVV: *rk = SK_VIR; *ck = SK_VIR; *ld = h->nv; return true;
        default: return false;
    }
}

int64_t lay_block_rows(const rcc_t* h, int kind, int b) {
    return (int64_t)h->sym.map[kind][b].size();
}

// offset of block b in the concatenated layout (b = nrep: total size)
int64_t lay_offset(const rcc_t* h, int rk, int ck, int b) {
    int64_t off = 0;
    for (int c = 0; c < b; ++c)
        off += (int64_t)h->sym.map[rk][c].size() * (int64_t)h->sym.map[ck][c].size();
    return off;
}

int ensure_stage(rcc_t* h, int64_t n) {
    if (n <= h->stage_n) return 0;
    if (h->stage) {
        auto it = std::find(h->allocs.begin(), h->allocs.end(), (void*)h->stage);
        if (it != h->allocs.end()) h->allocs.erase(it);
        HIPCHK(h, hipFree(h->stage));
        h->stage = nullptr;
        h->stage_n = 0;
    }
    CK(dalloc(h, &h->stage, n));
    h->stage_n = n;
    return 0;
}

// Whole symmetry-layout array (host or device) -> C1 array, zero elsewhere.
int sym_to_c1(rcc_t* h, int lay, const void* src, hipMemcpyKind kind, double2* dst,
              int64_t c1_elems) {
    int rk, ck;
    int64_t ld;
    lay_spec(h, lay, &rk, &ck, &ld);
    const SymInfo& S = h->sym;
    const int64_t total = lay_offset(h, rk, ck, S.nrep);
    const double2* dsrc = (const double2*)src;
    if (kind == hipMemcpyHostToDevice) {
        CK(ensure_stage(h, total));
        HIPCHK(h, hipMemcpyAsync(h->stage, src, (size_t)total * 16, kind, h->st));
        dsrc = h->stage;
    }
    k_fill_zero(dst, c1_elems, h->st);
    for (int b = 0; b < S.nrep; ++b)
        k_scatter_block(dsrc + lay_offset(h, rk, ck, b), lay_block_rows(h, rk, b),
                        lay_block_rows(h, ck, b), S.dmap + S.doff[rk][b], S.dmap + S.doff[ck][b],
                        dst, ld, h->st);
    HIPCHK(h, hipStreamSynchronize(h->st));
    return 0;
}

// C1 array <-> whole symmetry-layout array, both on the device.
BlockSet block_set(const rcc_t* h, int lay, int64_t* ld) {
    int rk, ck;
    lay_spec(h, lay, &rk, &ck, ld);
    const SymInfo& S = h->sym;
    BlockSet bs{};
    bs.nb = S.nrep;
    for (int b = 0; b < S.nrep; ++b) {
        bs.b[b].rmap = S.dmap + S.doff[rk][b];
        bs.b[b].cmap = S.dmap + S.doff[ck][b];
        bs.b[b].rows = lay_block_rows(h, rk, b);
        bs.b[b].cols = lay_block_rows(h, ck, b);
        bs.b[b].off = lay_offset(h, rk, ck, b);
    }
    return bs;
}

void c1_to_sym_dev(rcc_t* h, int lay, const double2* src, double2* dst) {
    int64_t ld;
    const BlockSet bs = block_set(h, lay, &ld);
    k_blockset(src, bs, dst, ld, true, h->st);
}

void sym_to_c1_dev(rcc_t* h, int lay, const double2* src, double2* dst) {
    int64_t ld;
    const BlockSet bs = block_set(h, lay, &ld);
    k_blockset(src, bs, dst, ld, false, h->st);
}

int64_t lay_total(const rcc_t* h, int lay) {
    int rk, ck;
    int64_t ld;
    lay_spec(h, lay, &rk, &ck, &ld);
    return lay_offset(h, rk, ck, h->sym.nrep);
}

// Descriptors of one grouped GEMM over the irrep blocks: C(b) = alpha A(b) op(B(b)) +
// beta C(b), each operand a symmetry-layout array (layouts la, lb, lc), block b of
// shape m(b) x k(b) etc. taken from the layouts; blocks with an empty side are left out.
int make_groups(rcc_t* h, int la, const double2* A, int lb, const double2* B, int lc,
                double2* C, bool bconj, double2 alpha, double2 beta, ZGroup** dG, int* ng,
                int64_t* nblocks, std::vector<ZGroup>* host) {
    std::vector<ZGroup> g;
    for (int b = 0; b < h->sym.nrep; ++b) {
        int64_t ar, ac, ao, br, bc, bo, cr, cc, co;
        rcc_layout_block(h, la, b, &ar, &ac, &ao);
        rcc_layout_block(h, lb, b, &br, &bc, &bo);
        rcc_layout_block(h, lc, b, &cr, &cc, &co);
        const int64_t k = ac;
        if (cr == 0 || cc == 0 || k == 0) continue;
        if (ar != cr || (bconj ? (bc != k || br != cc) : (br != k || bc != cc))) {
            h->err = "grouped GEMM: block shapes do not chain";
            return -1;
        }
        ZGroup z{};
        z.m = (int)cr;
        z.n = (int)cc;
        z.k = (int)k;
        z.A = A + ao;
        z.lda = std::max<int64_t>(ar, 1);
        z.B = B + bo;
        z.ldb = std::max<int64_t>(br, 1);
        z.C = C + co;
        z.ldc = std::max<int64_t>(cr, 1);
        z.alpha = alpha;
        z.beta = beta;
        g.push_back(z);
    }
    *ng = (int)g.size();
    *nblocks = zgroup_offsets(g.data(), *ng);
    *host = g;
    CK(dalloc_raw(h, dG, std::max<int64_t>(1, (int64_t)g.size()) * (int64_t)sizeof(ZGroup)));
    if (!g.empty())
        HIPCHK(h, hipMemcpy(*dG, g.data(), g.size() * sizeof(ZGroup), hipMemcpyHostToDevice));
    return 0;
}

int run_groups(rcc_t* h, int opb, int tag, const ZGroup* dG, int ng, int64_t nblocks) {
    if (g_record) {
        // recorded for a batched replay: the host copy of the same descriptors
        const int k = dG == h->g_ladder ? 0 : dG == h->g_x ? 1 : dG == h->g_ring1 ? 2 : 3;
        Rec r;
        r.kind = RK_GROUPED;
        r.groups = h->hg[k];
        r.opb = opb;
        r.tag = tag;
        r.gauss = h->gemm_gauss;
        g_record->push_back(std::move(r));
        return 0;
    }
    if (zgemm_grouped(opb, tag, h->gemm_gauss, dG, ng, nblocks, h->st) != 0) {
        h->err = "grouped GEMM launch failed";
        return -1;
    }
    return 0;
}

// The VVVV and VOVV irrep blocks of a symmetric handle (nrep > 1).
int alloc_blocks(rcc_t* h) {
    const SymInfo& S = h->sym;
    bool selfinv = true;
    for (int i = 0; i < S.nrep; ++i) selfinv = selfinv && S.ptab[i + S.nrep * i] == 0;
    h->ring_grouped = selfinv;
    CK(dalloc(h, &h->VVVV_s, lay_total(h, RCC_VVVV)));
    CK(dalloc(h, &h->VOVV_s, lay_total(h, RCC_VOVV)));
    k_fill_zero(h->VVVV_s, lay_total(h, RCC_VVVV), h->st);
    k_fill_zero(h->VOVV_s, lay_total(h, RCC_VOVV), h->st);
    HIPCHK(h, hipStreamSynchronize(h->st));
    return 0;
}

// Block copies and grouped-GEMM descriptors of the C1 working layout of a
// symmetric handle (after alloc_blocks and alloc_c1_work).
int setup_grouped(rcc_t* h) {
    CK(dalloc(h, &h->TAU_s, lay_total(h, RCC_LAY_T2)));
    CK(dalloc(h, &h->S2_s, lay_total(h, RCC_LAY_T2)));
    CK(dalloc(h, &h->X_s, lay_total(h, RCC_VOOO)));
    CK(make_groups(h, RCC_VVVV, h->VVVV_s, RCC_LAY_T2, h->TAU_s, RCC_LAY_T2, h->S2_s, false,
                   cplx(1, 0), cplx(1, 0), &h->g_ladder, &h->ng_ladder, &h->nb_ladder, &h->hg[0]));
    CK(make_groups(h, RCC_VOVV, h->VOVV_s, RCC_LAY_T2, h->TAU_s, RCC_VOOO, h->X_s, false,
                   cplx(1, 0), cplx(0, 0), &h->g_x, &h->ng_x, &h->nb_x, &h->hg[1]));
    if (h->ring_grouped) {
        const int64_t nr = lay_total(h, RCC_VOVO);
        CK(dalloc(h, &h->H_s, nr));
        CK(dalloc(h, &h->RING_s, nr));
        CK(dalloc(h, &h->RING2_s, nr));
        CK(dalloc(h, &h->VVOOR_s, nr));
        CK(make_groups(h, RCC_VOVO, h->RING2_s, RCC_VOVO, h->VVOOR_s, RCC_VOVO, h->H_s, true,
                       cplx(1, 0), cplx(1, 0), &h->g_ring1, &h->ng_ring1, &h->nb_ring1,
                       &h->hg[2]));
        CK(make_groups(h, RCC_VOVO, h->H_s, RCC_VOVO, h->RING_s, RCC_VOVO, h->RING2_s, false,
                       cplx(1, 0), cplx(0, 0), &h->g_ring2, &h->ng_ring2, &h->nb_ring2,
                       &h->hg[3]));
    }
    HIPCHK(h, hipStreamSynchronize(h->st));
    return 0;
}

// The C1 working set of a blocked handle, for a path that runs in the C1
// layout: allocated on first use, classes filled from the engine's blocks,
// T taken from the engine (c1_sync_t) before use.
int ensure_c1(rcc_t* h) {
    if (h->c1_ok) return 0;
    CK(alloc_c1_work(h));
    if (h->sym.nrep > 1) CK(setup_grouped(h));
    if (h->bk) {
        for (int c : {RCC_OOOO, RCC_VOOO, RCC_VVOO, RCC_VOVO, RCC_VOVV})
            sym_to_c1_dev(h, c, blk_ptr(h->bk, blk::A_OOOO + c), h->cls[c]);
        HIPCHK(h, hipStreamSynchronize(h->st));
        h->c1t_stale = true;
    }
    h->derived_ok = false;
    return 0;
}

// C1 array -> whole symmetry-layout array in host memory.
int c1_to_sym_host(rcc_t* h, int lay, const double2* src, double* out) {
    int rk, ck;
    int64_t ld;
    lay_spec(h, lay, &rk, &ck, &ld);
    const SymInfo& S = h->sym;
    const int64_t total = lay_offset(h, rk, ck, S.nrep);
    CK(ensure_stage(h, total));
    for (int b = 0; b < S.nrep; ++b)
        k_gather_block(src, ld, S.dmap + S.doff[rk][b], S.dmap + S.doff[ck][b],
                       lay_block_rows(h, rk, b), lay_block_rows(h, ck, b),
                       h->stage + lay_offset(h, rk, ck, b), h->st);
    HIPCHK(h, hipStreamSynchronize(h->st));
    HIPCHK(h, hipMemcpy(out, h->stage, (size_t)total * 16, hipMemcpyDeviceToHost));
    return 0;
}

// ------------------------------------------------------------------ Lambda
//
// relccsd's Lambda branches (CCEQN cceqns.F:277-288, T1EQNS :860-932 and the
// Lambda-only routines of ccintm.F) are switched off (ccdriv.F:1918-1924).
// The equations solved here are those of the CCSD Lagrangian
//     L(T, Lam) = E(T) + Lam . R(T),   R(T) = S(T) - D o T,
// with S the pre-DENOM right-hand side of one CCEQN pass (the term sequence
// above), D the DENOM denominators and E CCENG's complex sum zdotc(FVO, T1)
// + zdotc(VVOO, tau).  Integrals and Fock enter E and S conjugated or not,
// T never, so stationarity in T is holomorphic:
//     Lam = (dE/dT + Lam . dS/dT) / D,
// iterated like CCEQN (DENOM, ZCORE, DIISX).  Lam . dS/dT is the reverse
// (adjoint) sweep of the term sequence at the converged T: a contraction
// C += op(A) op(B) whose B depends on T sends Bbar += op(A)^T Cbar (and the
// same for A); each antisymmetrising pack turns into the matching expansion
// and back:
//     pack_pij^T = -expand_e2      expand_e2^T = pack_pij(-1)
//     pack_pab^T = expand_f1       expand_f1^T = pack_pab
//     pack_pabij^T = expand_ring   expand_ring^T = pack_pabij
//     perm_akci^T = perm_akci      trace_*^T = trace_*_adj.
// dE/dT1 = conj(FVO) + conj(VVOO_R) T1 = HOV (VVOO_R is symmetric under
// (ai) <-> (ck)) and dE/dT2 = conj(VVOO).

int lam_check(rcc_t* h) {
    if (h->precision != 0) {
        h->err = "Lambda equations: precision 0 (FP64) only";
        return -1;
    }
    return ensure_c1(h);
}

int lam_alloc(rcc_t* h) {
    if (h->lam_alloc) return 0;
    const int o = h->no, v = h->nv;
    const int64_t nvo = h->nvo, noot = h->noot, nvvt = h->nvvt, namp = h->nt1 + h->nt2;
    CK(dalloc(h, &h->LAM, namp));
    CK(dalloc(h, &h->LAMS, namp));
    CK(dalloc(h, &h->bTAU, h->nt2));
    CK(dalloc(h, &h->bX, nvo * noot));
    CK(dalloc(h, &h->bWC, nvo * noot));
    CK(dalloc(h, &h->bRING, nvo * nvo));
    CK(dalloc(h, &h->bRING2, nvo * nvo));
    CK(dalloc(h, &h->bH, nvo * nvo));
    CK(dalloc(h, &h->bE2, nvvt * o * o));
    CK(dalloc(h, &h->bF1, (int64_t)v * v * noot));
    CK(dalloc(h, &h->bP, nvo * std::max<int64_t>({noot, (int64_t)o, 1})));
    const int64_t nones = std::max<int64_t>({nvvt, noot, (int64_t)o, 1});
    CK(dalloc(h, &h->ones, nones));
    std::vector<double2> one((size_t)nones, cplx(1, 0));
    HIPCHK(h, hipMemcpy(h->ones, one.data(), (size_t)nones * 16, hipMemcpyHostToDevice));
    CK(dalloc(h, &h->bHO, (int64_t)o * o));
    CK(dalloc(h, &h->bHV, (int64_t)v * v));
    CK(dalloc(h, &h->bHOV, nvo));
    CK(dalloc(h, &h->bGO, (int64_t)o * o));
    CK(dalloc(h, &h->bGV, (int64_t)v * v));
    CK(dalloc(h, &h->bAZ, noot * noot));
    CK(dalloc(h, &h->bB2, (int64_t)o * o));
    CK(dalloc(h, &h->bW, nvo));
    CK(dalloc(h, &h->bFVOc, nvo));
    CK(dalloc(h, &h->LW, nvo));
    CK(dalloc(h, &h->LB2, (int64_t)o * o));
    CK(dalloc(h, &h->bPT2, nvvt + noot));
    if (h->sym.nrep > 1) {
        const int64_t nt2s = lay_total(h, RCC_LAY_T2);
        CK(dalloc(h, &h->S2b_s, nt2s));
        CK(dalloc(h, &h->bTAU_s, nt2s));
        // per compound irrep: (NVVT_c x NOOT_c) = VVVV_c^T (NVVT_c x NVVT_c) S2b_c
        std::vector<ZGroup> g;
        for (int b = 0; b < h->sym.nrep; ++b) {
            int64_t vr, vc, vo, tr, tc, to;
            rcc_layout_block(h, RCC_VVVV, b, &vr, &vc, &vo);
            rcc_layout_block(h, RCC_LAY_T2, b, &tr, &tc, &to);
            if (vr == 0 || tc == 0) continue;
            ZGroup z{};
            z.m = (int)vc;
            z.n = (int)tc;
            z.k = (int)vr;
            z.A = h->VVVV_s + vo;
            z.lda = std::max<int64_t>(vr, 1);
            z.B = h->S2b_s + to;
            z.ldb = std::max<int64_t>(tr, 1);
            z.C = h->bTAU_s + to;
            z.ldc = std::max<int64_t>(tr, 1);
            z.alpha = cplx(1, 0);
            z.beta = cplx(0, 0);
            g.push_back(z);
        }
        h->ng_ladder_adj = (int)g.size();
        h->nb_ladder_adj = zgroup_offsets(g.data(), h->ng_ladder_adj);
        CK(dalloc_raw(h, &h->g_ladder_adj, std::max<int64_t>(1, (int64_t)g.size()) *
                                                (int64_t)sizeof(ZGroup)));
        if (!g.empty())
            HIPCHK(h, hipMemcpy(h->g_ladder_adj, g.data(), g.size() * sizeof(ZGroup),
                                hipMemcpyHostToDevice));
    }
    h->lam_alloc = true;
    return 0;
}

// The forward quantities the sweep reuses, at the handle's current T: one
// pass of the term sequence (HO, HV, HOV, GO, GV, A, H, X, TAU, RING, WC;
// S is scratch here) and the T1EQNS products W, W^T T1.
int lam_setup(rcc_t* h) {
    CK(lam_check(h));
    CK(prep(h));
    CK(lam_alloc(h));
    CK(c1_sync_t(h));
    invalidate_t(h);
    CK(run_terms_c1(h));
    const int o = h->no, v = h->nv;
    k_conj(h->nvo, h->FVO, h->LW, h->st);
    k_axpby(h->nvo, cplx(1, 0), h->HOV, cplx(-2, 0), h->LW, h->st);
    return gemm(h, OP_T, OP_N, o, o, v, cplx(1, 0), h->LW, v, h->T1(), v, cplx(0, 0), h->LB2, o);
}

// out = [with_energy] dE/dT + Lin . dS/dT (packed T1 | T2 layouts).  The
// forward pass of lam_setup must be current.  Leaves the Fock-block adjoints
// bHO, bHV, bFVOc (d/d conj FVO, without the energy's T1) for the density.
int lam_sweep(rcc_t* h, const double2* Lin, double2* out, bool with_energy) {
    const int o = h->no, v = h->nv;
    const int64_t nvo = h->nvo, noot = h->noot, nvvt = h->nvvt, nt1 = h->nt1, nt2 = h->nt2;
    const int64_t K1 = (int64_t)v * noot, K2 = nvvt * o;
    hipStream_t st = h->st;
    const double2 one = cplx(1, 0), zero = cplx(0, 0), mone = cplx(-1, 0);
    const double2* s1b = Lin;
    const double2* s2b = Lin + nt1;
    double2* t1b = out;
    double2* t2b = out + nt1;
    const double2* T1 = h->T1();
    const double2* T2 = h->T2();
    if (with_energy) {
        HIPCHK(h, hipMemcpyAsync(t1b, h->HOV, (size_t)nt1 * 16, hipMemcpyDeviceToDevice, st));
        k_conj(nt2, h->cls[RCC_VVOO], t2b, st);
    } else {
        k_fill_zero(out, nt1 + nt2, st);
    }
    // ---- adjoint of the S2 terms
    // BINTM ladder S2 += VVVV TAU (ccintm.F:700-722):  TAUb = VVVV^T S2b (with
    // symmetry per compound irrep block; TAUb is zero off the blocks)
    ktick(h, 0, 0);
    if (h->sym.nrep > 1) {
        c1_to_sym_dev(h, RCC_LAY_T2, s2b, h->S2b_s);
        if (zgemm_grouped_ops(OP_T, OP_N, 0, h->gemm_gauss, h->g_ladder_adj, h->ng_ladder_adj,
                              h->nb_ladder_adj, st) != 0) {
            h->err = "grouped ladder adjoint launch failed";
            return -1;
        }
        k_fill_zero(h->bTAU, nt2, st);
        sym_to_c1_dev(h, RCC_LAY_T2, h->bTAU_s, h->bTAU);
    } else {
        CK(gemm(h, OP_T, OP_N, (int)nvvt, (int)noot, (int)nvvt, one, h->cls[RCC_VVVV], nvvt, s2b,
                nvvt, zero, h->bTAU, nvvt, 1, 0, 0, 0, 1));
    }
    ktick(h, 0, 1);
    // T2EQNS TAU A (cceqns.F:1013-1026):  TAUb += S2b A^T,  Ab = TAU^T S2b
    CK(gemm(h, OP_N, OP_T, (int)nvvt, (int)noot, (int)noot, one, s2b, nvvt, h->AZ, noot, one,
            h->bTAU, nvvt));
    CK(gemm(h, OP_T, OP_N, (int)noot, (int)noot, (int)nvvt, one, h->TAU, nvvt, s2b, nvvt, zero,
            h->bAZ, noot));
    // S2 += pack_pab(ZX) in BINTM and T2EQNS:  ZXb = expand_f1(S2b)
    double2* zxb = h->ZX;
    k_expand_f1(s2b, zxb, noot, h->vv, v, st);
    // BINTM ZX_b = -X_b T1^T (ccintm.F:760-829):  Xb_b = -ZXb_b T1
    CK(gemm(h, OP_N, OP_N, v, o, v, mone, zxb, v, T1, v, zero, h->bX, v, (int)noot,
            (int64_t)v * v, 0, nvo));
    // T1EQN2 (cceqns.F:2601-2714) S1 += trace(X):  Xb += trace^T(S1b)
    k_trace_t1eqn2_adj(s1b, h->bX, h->oo, v, o, st);
    // X = VOVV TAU (ccintm.F:772-788):  TAUb += VOVV^T Xb
    ktick(h, 3, 0);
    CK(gemm(h, OP_T, OP_N, (int)nvvt, (int)noot, (int)nvo, one, h->cls[RCC_VOVV], nvo, h->bX, nvo,
            one, h->bTAU, nvvt));
    ktick(h, 3, 1);
    // T1b -= sum_b (ZXb_b^T X_b + ZXb_b WC_b)  (BINTM; T2EQNS ZX_b -= T1 WC_b^T)
    CK(gemm(h, OP_T, OP_N, v, o, v, one, zxb, v, h->X, v, zero, h->bP, v, (int)noot,
            (int64_t)v * v, nvo, nvo));
    CK(gemm(h, OP_N, OP_N, v, o, v, one, zxb, v, h->WC, v, one, h->bP, v, (int)noot,
            (int64_t)v * v, nvo, nvo));
    CK(gemm(h, OP_N, OP_N, (int)nvo, 1, (int)noot, mone, h->bP, nvo, h->ones, noot, one, t1b,
            nvo));
    // WCb_b = -ZXb_b^T T1;  WC = -VOOO + pack_pij(Y1, -1):  Y1b = expand_e2(WCb);
    // Y1_i = VOVO_i T1 (cceqns.F:1158-1221):  T1b += sum_i VOVO_i^T Y1b_i
    if (o > 1) {
        CK(gemm(h, OP_T, OP_N, v, o, v, mone, zxb, v, T1, v, zero, h->bWC, v, (int)noot,
                (int64_t)v * v, 0, nvo));
        double2* y1b = h->Y1;
        k_expand_e2(h->bWC, y1b, nvo, h->oo, o, st);
        CK(gemm(h, OP_T, OP_N, v, o, (int)nvo, one, h->cls[RCC_VOVO], nvo, y1b, nvo, zero, h->bP,
                v, o, (int64_t)v * nvo, nvo * o, nvo));
        CK(gemm(h, OP_N, OP_N, (int)nvo, 1, o, one, h->bP, nvo, h->ones, o, one, t1b, nvo));
    }
    // T2EQNS GV F1(T2) (cceqns.F:1183-1184):  GVb = ZXb F1^T,  F1b = GV^T ZXb,
    // T2b += pack_pab(F1b)
    k_expand_f1(T2, h->F1, noot, h->vv, v, st);
    CK(gemm(h, OP_N, OP_T, v, v, (int)K1, one, zxb, v, h->F1, v, zero, h->bGV, v));
    CK(gemm(h, OP_T, OP_N, v, (int)K1, v, one, h->GV, v, zxb, v, zero, h->bF1, v));
    k_pack_pab(h->bF1, t2b, noot, 1.0, h->vv, v, st);
    // T2EQNS S2 += pack_pij(YP), YP_i = VOVV_i^H T1 + E2(T2) GO (cceqns.F:1051-1139):
    // YPb = -expand_e2(S2b) (the sign is carried by the products' alpha)
    if (o > 1) {
        double2* ypb = h->YP;
        k_expand_e2(s2b, ypb, nvvt, h->oo, o, st);
        CK(gemm(h, OP_R, OP_N, v, o, (int)nvvt, mone, h->cls[RCC_VOVV], nvo, ypb, nvvt, zero,
                h->bP, v, o, v, nvvt * o, nvo));
        CK(gemm(h, OP_N, OP_N, (int)nvo, 1, o, one, h->bP, nvo, h->ones, o, one, t1b, nvo));
        k_expand_e2(T2, h->E2, nvvt, h->oo, o, st);
        CK(gemm(h, OP_N, OP_T, (int)K2, o, o, mone, ypb, K2, h->GO, o, zero, h->bE2, K2));
        CK(gemm(h, OP_T, OP_N, o, o, (int)K2, mone, h->E2, K2, ypb, K2, zero, h->bGO, o));
        k_pack_pij(h->bE2, t2b, nvvt, -1.0, h->oo, o, st);
    } else {
        k_fill_zero(h->bGO, (int64_t)o * o, st);
    }
    // T2EQNS ring S2 += pack_pabij(H RING) (cceqns.F:1243-1263):
    // RING2b = expand_ring(S2b),  Hb = RING2b RING^T,  RINGb = H^T RING2b
    k_expand_ring(s2b, nullptr, 1.0, h->bRING2, h->vv, h->oo, v, o, st);
    ktick(h, 2, 0);
    CK(gemm(h, OP_N, OP_T, (int)nvo, (int)nvo, (int)nvo, one, h->bRING2, nvo, h->RING, nvo, zero,
            h->bH, nvo, 1, 0, 0, 0, 2));
    CK(gemm(h, OP_T, OP_N, (int)nvo, (int)nvo, (int)nvo, one, h->H, nvo, h->bRING2, nvo, zero,
            h->bRING, nvo, 1, 0, 0, 0, 2));
    ktick(h, 2, 1);
    // ---- adjoint of T1EQNS (cceqns.F:645-935)
    // RING HOV:  RINGb += S1b HOV^T,  HOVb = RING^T S1b;  T2b += pack_pabij(RINGb)
    CK(gemm(h, OP_N, OP_T, (int)nvo, (int)nvo, 1, one, s1b, nvo, h->HOV, nvo, one, h->bRING,
            nvo));
    CK(gemm(h, OP_T, OP_N, (int)nvo, 1, (int)nvo, one, h->RING, nvo, s1b, nvo, zero, h->bHOV,
            nvo));
    k_pack_pabij(h->bRING, t2b, h->vv, h->oo, v, o, st, 1.0);
    // T1 (W^T T1), W = HOV - 2 conj(FVO):  T1b += S1b B2^T + W B2b,  B2b = T1^T S1b,
    // Wb = T1 B2b^T;  HOVb += Wb,  d/d conj(FVO) = -2 Wb
    CK(gemm(h, OP_N, OP_T, v, o, o, one, s1b, v, h->LB2, o, one, t1b, v));
    CK(gemm(h, OP_T, OP_N, o, o, v, one, T1, v, s1b, v, zero, h->bB2, o));
    CK(gemm(h, OP_N, OP_T, v, o, o, one, T1, v, h->bB2, o, zero, h->bW, v));
    CK(gemm(h, OP_N, OP_N, v, o, o, one, h->LW, v, h->bB2, o, one, t1b, v));
    k_axpby(nvo, one, h->bW, one, h->bHOV, st);
    k_axpby(nvo, cplx(-2, 0), h->bW, zero, h->bFVOc, st);
    // HV T1 - T1 HO:  HVb = S1b T1^T,  HOb = -T1^T S1b,  T1b += HV^T S1b - S1b HO^T
    CK(gemm(h, OP_N, OP_T, v, v, o, one, s1b, v, T1, v, zero, h->bHV, v));
    CK(gemm(h, OP_T, OP_N, v, o, v, one, h->HV, v, s1b, v, one, t1b, v));
    CK(gemm(h, OP_N, OP_T, v, o, o, mone, s1b, v, h->HO, o, one, t1b, v));
    CK(gemm(h, OP_T, OP_N, o, o, v, mone, T1, v, s1b, v, zero, h->bHO, o));
    // - VOVO_R T1:  T1b -= VOVO_R^T S1b
    CK(gemm(h, OP_T, OP_N, (int)nvo, 1, (int)nvo, mone, h->VOVO_R, nvo, s1b, nvo, one, t1b,
            nvo));
    // T1EQN1 F1(TAU) M1 (cceqns.F:2521-2599):  F1b = S1b M1^T  (HVINTM adds to it below)
    CK(gemm(h, OP_N, OP_T, v, (int)K1, o, one, s1b, v, h->M1, K1, zero, h->bF1, v));
    // ---- adjoint of HINTM (ccintm.F:221-453): H = perm(HAKCI) + RING2' VVOO_R^H,
    // RING2' = expand_ring(T2, T1, 1/2):  RING2'b = Hb conj(VVOO_R)
    ktick(h, 1, 0);
    CK(gemm(h, OP_N, OP_R, (int)nvo, (int)nvo, (int)nvo, one, h->bH, nvo, h->VVOO_R, nvo, zero,
            h->bRING2, nvo, 1, 0, 0, 0, 2));
    ktick(h, 1, 1);
    // T2b += pack_pabij(RING2'b, 1/2);  the -T1(a,k) T1(c,i) part through
    // Q[(a,k),(c,i)] = RING2'b[(a,i),(c,k)]:  T1b -= Q T1 + Q^T T1
    k_pack_pabij(h->bRING2, t2b, h->vv, h->oo, v, o, st, 0.5);
    k_perm_akci(h->bRING2, h->RING2, v, o, false, st);
    CK(gemm(h, OP_N, OP_N, (int)nvo, 1, (int)nvo, mone, h->RING2, nvo, T1, nvo, one, t1b, nvo));
    CK(gemm(h, OP_T, OP_N, (int)nvo, 1, (int)nvo, mone, h->RING2, nvo, T1, nvo, one, t1b, nvo));
    // HAKCIb = perm(Hb);  HAKCI += T1 MH (ccintm.F:268-273):  T1b += HAKCIb MH^T
    double2* hakb = h->HAKCI;
    k_perm_akci(h->bH, hakb, v, o, false, st);
    CK(gemm(h, OP_N, OP_T, v, o, (int)(o * nvo), one, hakb, v, h->MH, o, one, t1b, v));
    // GVINTM's trace of HAKCI_gv (ccintm.F:208-218):  HAKCIb -= trace^T(GVb);
    // HAKCI_gv = -VOVV_Q T1:  T1b -= VOVV_Q^T HAKCIb
    k_trace_gv_adj(h->bGV, hakb, -1.0, v, o, st);
    CK(gemm(h, OP_T, OP_N, v, o, (int)(nvo * v), mone, h->VOVV_Q, nvo * v, hakb, nvo * v, one, t1b,
            v));
    // GVINTM GV = HV - T1 FVO^H:  HVb += GVb,  T1b -= GVb conj(FVO),  FVOcb -= GVb^T T1
    k_axpby((int64_t)v * v, one, h->bGV, one, h->bHV, st);
    CK(gemm(h, OP_N, OP_R, v, o, v, mone, h->bGV, v, h->FVO, v, one, t1b, v));
    CK(gemm(h, OP_T, OP_N, v, o, v, mone, h->bGV, v, T1, v, one, h->bFVOc, v));
    // GOINTM GO = HO + FVO^H T1 - MGO^T T1 (ccintm.F:95-146)
    k_axpby((int64_t)o * o, one, h->bGO, one, h->bHO, st);
    CK(gemm(h, OP_R, OP_N, v, o, o, one, h->FVO, v, h->bGO, o, one, t1b, v));
    CK(gemm(h, OP_N, OP_T, v, o, o, one, T1, v, h->bGO, o, one, h->bFVOc, v));
    CK(gemm(h, OP_N, OP_N, (int)nvo, 1, o * o, mone, h->MGO, nvo, h->bGO, (int64_t)o * o, one,
            t1b, nvo));
    // AINTM A = OOOO + pack_pij(YAZ) + VVOO^H TAU (ccintm.F:34-93):
    // TAUb += conj(VVOO) Ab;  YAZb = -expand_e2(Ab);  T1b += sum_i conj(VOOO_i) YAZb_i
    CK(gemm(h, OP_R, OP_N, (int)nvvt, (int)noot, (int)noot, one, h->cls[RCC_VVOO], nvvt, h->bAZ,
            noot, one, h->bTAU, nvvt));
    if (o > 0 && noot > 0) {
        double2* yazb = h->YAZ;
        k_expand_e2(h->bAZ, yazb, noot, h->oo, o, st);
        CK(gemm(h, OP_R, OP_N, v, o, (int)noot, mone, h->cls[RCC_VOOO], nvo, yazb, noot, zero,
                h->bP, v, o, v, noot * o, nvo));
        CK(gemm(h, OP_N, OP_N, (int)nvo, 1, o, one, h->bP, nvo, h->ones, o, one, t1b, nvo));
    }
    // HOVINTM HOV = conj(FVO) + VVOO_R^H T1 (ccintm.F:514-565):
    // FVOcb += HOVb,  T1b += conj(VVOO_R) HOVb
    k_axpby(nvo, one, h->bHOV, one, h->bFVOc, st);
    CK(gemm(h, OP_R, OP_N, (int)nvo, 1, (int)nvo, one, h->VVOO_R, nvo, h->bHOV, nvo, one, t1b,
            nvo));
    // HVINTM HV = FVV' - F1(TAU) VVOO_F1^H (ccintm.F:567-623):
    // F1b -= HVb conj(VVOO_F1);  TAUb += pack_pab(F1b)
    CK(gemm(h, OP_N, OP_R, v, (int)K1, v, mone, h->bHV, v, h->VVOO_F1, v, one, h->bF1, v));
    k_pack_pab(h->bF1, h->bTAU, noot, 1.0, h->vv, v, st);
    // HOINTM HO = FOO' + VVOO_E2^H E2(TAU) (ccintm.F:455-512):
    // E2b = conj(VVOO_E2) HOb;  TAUb += pack_pij(E2b, -1)
    CK(gemm(h, OP_R, OP_N, (int)K2, o, o, one, h->VVOO_E2, K2, h->bHO, o, zero, h->bE2, K2));
    k_pack_pij(h->bE2, h->bTAU, nvvt, -1.0, h->oo, o, st);
    // GETTAU (ccintm.F:833-928):  T2b += TAUb,  T1b += expand_ring(TAUb) T1
    k_axpby(nt2, one, h->bTAU, one, t2b, st);
    k_expand_ring(h->bTAU, nullptr, 1.0, h->RING2, h->vv, h->oo, v, o, st);
    return gemm(h, OP_N, OP_N, (int)nvo, 1, (int)nvo, one, h->RING2, nvo, T1, nvo, one, t1b, nvo);
}

void reset_diis(rcc_t* h) {
    h->it = 0;
    h->itc = 0;
    std::fill(h->bb2.begin(), h->bb2.end(), 0.0);
}

// Lam at the current T, from Lam0 = (dE/dT) / D, MAXIT / NTOL / MAXDIM of the
// handle.  The DIIS history is the Lambda one afterwards, so it is cleared
// for a later CCEQN.
int lam_solve(rcc_t* h, int* nit, double* rms_out) {
    CK(lam_setup(h));
    CK(alloc_diis(h, h->nt1 + h->nt2));
    const int64_t nt1 = h->nt1, nt2 = h->nt2;
    hipStream_t st = h->st;
    const int* frz = h->any_frozen ? h->frozen : nullptr;
    HIPCHK(h, hipMemcpyAsync(h->LAM, h->HOV, (size_t)nt1 * 16, hipMemcpyDeviceToDevice, st));
    k_conj(nt2, h->cls[RCC_VVOO], h->LAM + nt1, st);
    k_denom(h->LAM, h->LAM + nt1, h->eps, frz, h->vv, h->oo, h->nv, h->no, st);
    h->lam_set = true;
    reset_diis(h);
    const double conv = std::pow(10.0, -h->ntol);
    double r = 1.0;
    int n = 0;
    while (n < h->maxit) {
        HIPCHK(h, hipEventRecord(h->ev0, st));
        CK(lam_sweep(h, h->LAM, h->LAMS, true));
        k_denom(h->LAMS, h->LAMS + nt1, h->eps, frz, h->vv, h->oo, h->nv, h->no, st);
        CK(diis_begin(h, h->LAM, h->LAMS, nt1 + nt2));
        HIPCHK(h, hipStreamSynchronize(st));
        CK(diis_end(h, &r, h->LAMS, nt1 + nt2));
        HIPCHK(h, hipMemcpyAsync(h->LAM, h->LAMS, (size_t)(nt1 + nt2) * 16,
                                 hipMemcpyDeviceToDevice, st));
        HIPCHK(h, hipEventRecord(h->ev1, st));
        HIPCHK(h, hipStreamSynchronize(st));
        float ms = 0.f;
        HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->last_lambda_ms = ms;
        for (int w = 0; w < 4; ++w) {
            float km = 0.f;
            if (h->ktimed[w] && hipEventElapsedTime(&km, h->kev[w][0], h->kev[w][1]) == hipSuccess)
                h->kms[w] = km;
            h->ktimed[w] = false;
        }
        ++n;
        if (r < conv) break;
    }
    reset_diis(h);
    if (nit) *nit = n;
    if (rms_out) *rms_out = r;
    return 0;
}

// The response density at (T, LAM): D (n x n, occupied then virtual,
// column-major, Hermitian) with d Re(Ec) / d lam = tr(D V) for f -> f + lam V.
// From the Fock-block adjoints of one sweep: dL/dFOO = HOb (off-diagonal;
// HOINTM zeroes the diagonal, ccintm.F:495), dL/dFVV = HVb (off-diagonal,
// ccintm.F:607), dL/dFVO = Lam1 (T1EQNS starts from FVO, cceqns.F:716),
// dL/d conj(FVO) = FVOcb + T1 (CCENG's zdotc), and dL/d eps from D o T:
//   eps_i: -sum Lam1 T1 over (a, i) - sum Lam2 T2 over the pairs holding i,
//   eps_a: +sum Lam1 T1 over (a, i) + sum Lam2 T2 over the pairs holding a.
// For Hermitian V the real part of dL/dlam is tr(D V) with
//   D_oo = (conj(HOb) + HOb^T) / 2, D_vv likewise, D_vo = (conj(Lam1) +
//   FVOcb + T1) / 2, D_ov = D_vo^H and the diagonal Re(dL/d eps).
int lam_density(rcc_t* h, double* dens) {
    CK(lam_setup(h));
    CK(lam_sweep(h, h->LAM, h->LAMS, true));
    const int o = h->no, v = h->nv, n = o + v;
    const int64_t nvo = h->nvo, noot = h->noot, nvvt = h->nvvt, nt1 = h->nt1, nt2 = h->nt2;
    hipStream_t st = h->st;
    double2* P = h->LAMS;       // Lam o T (the sweep's output is not needed here)
    k_hadamard(nt1 + nt2, h->LAM, h->T, P, st);
    // pair sums of Lam2 o T2: r over the IJ columns of each AB row, c over the AB rows
    CK(gemm(h, OP_N, OP_N, (int)nvvt, 1, (int)noot, cplx(1, 0), P + nt1, nvvt, h->ones, noot,
            cplx(0, 0), h->bPT2, nvvt));
    CK(gemm(h, OP_T, OP_N, (int)noot, 1, (int)nvvt, cplx(1, 0), P + nt1, nvvt, h->ones, nvvt,
            cplx(0, 0), h->bPT2 + nvvt, noot));
    std::vector<double2> p1(nvo), rs(nvvt), cs(noot), ho((size_t)o * o), hv((size_t)v * v),
        fvoc(nvo), l1(nvo), t1(nvo);
    auto d2h = [&](std::vector<double2>& dst, const double2* src) -> int {
        if (!dst.empty())
            HIPCHK(h, hipMemcpyAsync(dst.data(), src, dst.size() * 16, hipMemcpyDeviceToHost, st));
        return 0;
    };
    CK(d2h(p1, P));
    CK(d2h(rs, h->bPT2));
    CK(d2h(cs, h->bPT2 + nvvt));
    CK(d2h(ho, h->bHO));
    CK(d2h(hv, h->bHV));
    CK(d2h(fvoc, h->bFVOc));
    CK(d2h(l1, h->LAM));
    CK(d2h(t1, h->T1()));
    HIPCHK(h, hipStreamSynchronize(st));
    std::vector<double> epsb((size_t)n, 0.0);
    for (int i = 0; i < o; ++i)
        for (int a = 0; a < v; ++a) {
            const double x = p1[a + (size_t)v * i].x;
            epsb[i] -= x;
            epsb[o + a] += x;
        }
    for (int i = 0; i < o; ++i)
        for (int j = 0; j < o; ++j)
            if (i != j) epsb[i] -= cs[i > j ? pair_index(i, j, o) : pair_index(j, i, o)].x;
    for (int a = 0; a < v; ++a)
        for (int b = 0; b < v; ++b)
            if (a != b) epsb[o + a] += rs[a > b ? pair_index(a, b, v) : pair_index(b, a, v)].x;
    auto D = [&](int p, int q) -> double* { return dens + 2 * ((size_t)p + (size_t)n * q); };
    auto put = [&](int p, int q, double re, double im) {
        D(p, q)[0] = re;
        D(p, q)[1] = im;
    };
    for (int i = 0; i < o; ++i)
        for (int k = 0; k < o; ++k) {
            if (i == k) { put(k, k, epsb[k], 0.0); continue; }
            const double2 a = ho[k + (size_t)o * i], b = ho[i + (size_t)o * k];
            put(k, i, 0.5 * (a.x + b.x), 0.5 * (-a.y + b.y));
        }
    for (int c = 0; c < v; ++c)
        for (int a = 0; a < v; ++a) {
            if (a == c) { put(o + a, o + a, epsb[o + a], 0.0); continue; }
            const double2 x = hv[a + (size_t)v * c], y = hv[c + (size_t)v * a];
            put(o + a, o + c, 0.5 * (x.x + y.x), 0.5 * (-x.y + y.y));
        }
    for (int i = 0; i < o; ++i)
        for (int a = 0; a < v; ++a) {
            const size_t e = a + (size_t)v * i;
            const double re = 0.5 * (l1[e].x + fvoc[e].x + t1[e].x);
            const double im = 0.5 * (-l1[e].y + fvoc[e].y + t1[e].y);
            put(o + a, i, re, im);
            put(i, o + a, re, -im);
        }
    return 0;
}


// ---------------------------------------------------- grouped batch iteration
// Many independent handles: the four dominant products of every handle (X =
// VOVV.tau, the HINTM and T2EQNS rings, the ladder; per irrep block for a
// symmetric handle, one block for a C1 one) run as ONE grouped launch per
// term over all handles, on a batch stream that joins the handles' streams
// with events; the rest of each handle's term sequence stays on its own
// stream.  A plan (descriptors on the device, streams, events) is built once
// per list of handles.
struct BatchPlan {
    std::vector<rcc_t*> hs;
    ZGroup* dG[4] = {nullptr, nullptr, nullptr, nullptr};
    int ng[4] = {0, 0, 0, 0};
    int64_t nb[4] = {0, 0, 0, 0};
    int gauss = 1;
    hipStream_t st = nullptr;
    hipEvent_t evb = nullptr;
    std::vector<hipEvent_t> evh;
    // timing events of a compiled batch iteration (start and end on st)
    hipEvent_t et0 = nullptr, et1 = nullptr;
    // the whole term sequence of every handle, recorded once and compiled to
    // one launch per position over all handles (null: handles differ in
    // structure, the segmented form above runs instead); optionally replayed
    // from a hipGraph
    BatchProgram* prog = nullptr;
    std::string prog_err;
    int runs = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // batched DIISX / CCENG: parameters (pinned host staging and device copy)
    // and every handle's reduction results gathered for one transfer
    unsigned char* fin_host = nullptr;
    void* fin_dev = nullptr;
    size_t fin_bytes = 0;
    double2* dslots = nullptr;
    double2* hslots = nullptr;
};
std::vector<BatchPlan*> g_plans;

void free_plan(BatchPlan* p) {
    if (p->st) (void)hipStreamSynchronize(p->st);
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    batch_free(p->prog);
    if (p->fin_host) (void)hipHostFree(p->fin_host);
    if (p->fin_dev) (void)hipFree(p->fin_dev);
    if (p->dslots) (void)hipFree(p->dslots);
    if (p->hslots) (void)hipHostFree(p->hslots);
    for (int k = 0; k < 4; ++k)
        if (p->dG[k]) (void)hipFree(p->dG[k]);
    for (hipEvent_t e : p->evh) (void)hipEventDestroy(e);
    if (p->evb) (void)hipEventDestroy(p->evb);
    if (p->et0) (void)hipEventDestroy(p->et0);
    if (p->et1) (void)hipEventDestroy(p->et1);
    if (p->st) (void)hipStreamDestroy(p->st);
    delete p;
}

void drop_plans_of(rcc_t* h) {
    for (size_t i = 0; i < g_plans.size();) {
        BatchPlan* p = g_plans[i];
        if (std::find(p->hs.begin(), p->hs.end(), h) != p->hs.end()) {
            free_plan(p);
            g_plans.erase(g_plans.begin() + i);
        } else {
            ++i;
        }
    }
}

// The descriptors of a C1 FP64 handle's four products (one group each).
std::vector<ZGroup> c1_group(rcc_t* h, int kind) {
    ZGroup z{};
    const int64_t nvo = h->nvo, noot = h->noot, nvvt = h->nvvt;
    z.alpha = cplx(1, 0);
    if (kind == 0) {          // S2 += VVVV TAU
        z.m = (int)nvvt; z.n = (int)noot; z.k = (int)nvvt;
        z.A = h->cls[RCC_VVVV]; z.lda = nvvt; z.B = h->TAU; z.ldb = nvvt;
        z.C = h->S2(); z.ldc = nvvt; z.beta = cplx(1, 0);
    } else if (kind == 1) {   // X = VOVV TAU
        z.m = (int)nvo; z.n = (int)noot; z.k = (int)nvvt;
        z.A = h->cls[RCC_VOVV]; z.lda = nvo; z.B = h->TAU; z.ldb = nvvt;
        z.C = h->X; z.ldc = nvo; z.beta = cplx(0, 0);
    } else if (kind == 2) {   // H += RING2' VVOO_R^H
        z.m = (int)nvo; z.n = (int)nvo; z.k = (int)nvo;
        z.A = h->RING2; z.lda = nvo; z.B = h->VVOO_R; z.ldb = nvo;
        z.C = h->H; z.ldc = nvo; z.beta = cplx(1, 0);
    } else {                  // RING2 = H RING
        z.m = (int)nvo; z.n = (int)nvo; z.k = (int)nvo;
        z.A = h->H; z.lda = nvo; z.B = h->RING; z.ldb = nvo;
        z.C = h->RING2; z.ldc = nvo; z.beta = cplx(0, 0);
    }
    z.lda = std::max<int64_t>(z.lda, 1);
    z.ldb = std::max<int64_t>(z.ldb, 1);
    z.ldc = std::max<int64_t>(z.ldc, 1);
    if (z.m <= 0 || z.n <= 0 || z.k <= 0) return {};
    return {z};
}

bool batch_groupable(rcc_t* const* hs, int n) {
    for (int i = 0; i < n; ++i) {
        rcc_t* h = hs[i];
        if (h->precision != 0 || h->gemm_gauss != hs[0]->gemm_gauss) return false;
        if (!h->bk && h->sym.nrep > 1 && !h->ring_grouped) return false;
    }
    return n > 0;
}

int get_plan(rcc_t* const* hs, int n, BatchPlan** out) {
    std::vector<rcc_t*> key(hs, hs + n);
    for (BatchPlan* p : g_plans)
        if (p->hs == key) {
            *out = p;
            return 0;
        }
    auto* p = new BatchPlan();
    p->hs = key;
    p->gauss = hs[0]->gemm_gauss;
    rcc_t* h0 = hs[0];
    auto fail = [&](const std::string& m) {
        h0->err = m;
        free_plan(p);
        return -1;
    };
    if (hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&p->evb, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&p->et0) != hipSuccess || hipEventCreate(&p->et1) != hipSuccess)
        return fail("batch plan: stream / event creation failed");
    for (int i = 0; i < n; ++i) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return fail("batch plan: event creation failed");
        p->evh.push_back(e);
    }
    for (int k = 0; k < 4; ++k) {
        std::vector<ZGroup> all;
        for (int i = 0; i < n; ++i) {
            const std::vector<ZGroup> g = hs[i]->sym.nrep > 1 ? hs[i]->hg[k] : c1_group(hs[i], k);
            all.insert(all.end(), g.begin(), g.end());
        }
        p->ng[k] = (int)all.size();
        p->nb[k] = zgroup_offsets(all.data(), p->ng[k]);
        if (all.empty()) continue;
        if (hipMalloc((void**)&p->dG[k], all.size() * sizeof(ZGroup)) != hipSuccess ||
            hipMemcpy(p->dG[k], all.data(), all.size() * sizeof(ZGroup), hipMemcpyHostToDevice) !=
                hipSuccess)
            return fail("batch plan: descriptor upload failed");
    }
    // record every handle's term sequence (nothing is launched) and compile it
    std::vector<std::vector<Rec>> recs((size_t)n);
    for (int i = 0; i < n; ++i) {
        rcc_t* h = hs[i];
        if ((h->bk ? bprep(h) : prep(h)) != 0) return fail(h->err);
        HIPCHK(h, hipStreamSynchronize(h->st));
        invalidate_t(h);
        h->capturing = true;
        g_record = &recs[i];
        const int rc = run_terms(h);
        g_record = nullptr;
        h->capturing = false;
        invalidate_t(h);
        if (rc != 0) return fail("batch plan: recording failed: " + h->err);
    }
    std::vector<std::vector<Rec>*> ptrs;
    for (auto& r : recs) ptrs.push_back(&r);
    if (batch_compile(ptrs, &p->prog, p->prog_err) != 0) p->prog = nullptr;
    if (p->prog && zgemm_autotune_enabled()) {
        if (batch_tune(p->prog, p->st) != 0) return fail("batch plan: tile tuning failed");
        HIPCHK(h0, hipStreamSynchronize(p->st));
    }
    g_plans.push_back(p);
    *out = p;
    return 0;
}

// The compiled term sequence of all handles on the batch stream, joined to
// the handles' streams; with use_graph on every handle the sequence is
// captured once (second call) and replayed from the graph afterwards.
// Everything up to the end of batch_finish runs on the batch stream, which
// the host synchronises before returning: the handles' streams need no join
// afterwards, and before the phase only a stream with work still queued does
// (two HIP calls per such handle instead of five per handle per iteration).
int batch_program_phase(BatchPlan* p, bool use_graph) {
    rcc_t* h0 = p->hs[0];
    for (size_t i = 0; i < p->hs.size(); ++i) {
        rcc_t* h = p->hs[i];
        const hipError_t q = hipStreamQuery(h->st);
        if (q == hipSuccess) continue;
        (void)hipGetLastError();      // "not ready" is a state, not an error
        if (q != hipErrorNotReady) {
            h->err = std::string("hipStreamQuery on the handle's stream: ") + hipGetErrorString(q);
            return -1;
        }
        HIPCHK(h0, hipEventRecord(p->evh[i], h->st));
        HIPCHK(h0, hipStreamWaitEvent(p->st, p->evh[i], 0));
    }
    HIPCHK(h0, hipEventRecord(p->et0, p->st));
    p->runs += 1;
    if (use_graph && p->gexec) {
        HIPCHK(h0, hipGraphLaunch(p->gexec, p->st));
    } else if (use_graph && p->runs >= 2) {
        HIPCHK(h0, hipStreamBeginCapture(p->st, hipStreamCaptureModeThreadLocal));
        const int rc = batch_run(p->prog, p->st);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(p->st, &g);
        if (rc != 0 || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            h0->err = "batch: graph capture of the term sequence failed";
            return -1;
        }
        p->graph = g;
        HIPCHK(h0, hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0));
        HIPCHK(h0, hipGraphLaunch(p->gexec, p->st));
    } else if (batch_run(p->prog, p->st) != 0) {
        h0->err = "batch: replay launch failed";
        return -1;
    }
    return 0;
}
// DIISX, CCENG and S -> T of every handle as batched launches on the batch
// stream, in the order finish_a / finish_b run them per handle: error vector
// and store, |d|^2 and sum d, the B-row dots (one gather and one transfer of
// all handles' results), the host FLIN per handle, then the extrapolation,
// tau(S), the two energy dots and S -> T (one more gather and transfer).
// Same kernels and partitioning as the per-handle path: same results.
struct FinBuilder {
    std::vector<unsigned char> host;
    struct Seg { int kind; size_t off; int64_t gx; };
    std::vector<Seg> segs;
    // appends one parameter block per handle of `kind`
    void add(int kind, const std::vector<std::vector<unsigned char>>& per, int64_t gx) {
        const size_t ps = batch_param_size(kind);
        const size_t off = (host.size() + 255) / 256 * 256;
        host.resize(off + ps * per.size());
        for (size_t i = 0; i < per.size(); ++i) memcpy(host.data() + off + i * ps, per[i].data(), ps);
        segs.push_back({kind, off, gx});
    }
};

int fin_run(BatchPlan* p, FinBuilder& fb, size_t slot_elems) {
    rcc_t* h0 = p->hs[0];
    const int nh = (int)p->hs.size();
    if (fb.host.size() > p->fin_bytes) {
        if (p->fin_host) (void)hipHostFree(p->fin_host);
        if (p->fin_dev) (void)hipFree(p->fin_dev);
        p->fin_host = nullptr;
        p->fin_dev = nullptr;
        p->fin_bytes = fb.host.size() * 2;
        HIPCHK(h0, hipHostMalloc((void**)&p->fin_host, p->fin_bytes, hipHostMallocDefault));
        HIPCHK(h0, hipMalloc(&p->fin_dev, p->fin_bytes));
    }
    memcpy(p->fin_host, fb.host.data(), fb.host.size());
    HIPCHK(h0, hipMemcpyAsync(p->fin_dev, p->fin_host, fb.host.size(), hipMemcpyHostToDevice,
                              p->st));
    for (const FinBuilder::Seg& sg : fb.segs)
        if (k_many(sg.kind, (const unsigned char*)p->fin_dev + sg.off, nh, sg.gx, p->st) != 0) {
            h0->err = "batched DIIS / energy launch failed";
            return -1;
        }
    HIPCHK(h0, hipMemcpyAsync(p->hslots, p->dslots, slot_elems * sizeof(double2),
                              hipMemcpyDeviceToHost, p->st));
    return 0;
}

template <class F>
std::vector<std::vector<unsigned char>> per_handle(BatchPlan* p, int kind, F fill) {
    std::vector<std::vector<unsigned char>> v(p->hs.size(),
                                              std::vector<unsigned char>(batch_param_size(kind)));
    for (size_t i = 0; i < p->hs.size(); ++i) fill(p->hs[i], (int)i, (void*)v[i].data());
    return v;
}

int batch_finish(BatchPlan* p, int* nit, double* ecc, double* rms) {
    rcc_t* h0 = p->hs[0];
    const int nh = (int)p->hs.size();
    const size_t slots = (size_t)nh * RED_SLOTS;
    if (!p->dslots) {
        HIPCHK(h0, hipMalloc((void**)&p->dslots, slots * sizeof(double2)));
        HIPCHK(h0, hipHostMalloc((void**)&p->hslots, slots * sizeof(double2), hipHostMallocDefault));
    }
    std::vector<int> cpos(nh);
    int64_t nmax = 0, itmax = 0;
    int bmax = 1;
    for (int i = 0; i < nh; ++i) {
        rcc_t* h = p->hs[i];
        cpos[i] = diis_advance(h);
        nmax = std::max<int64_t>(nmax, amp_n(h));
        itmax = std::max<int64_t>(itmax, h->it);
        bmax = std::max(bmax, red_blocks(amp_n(h), RED_BLOCKS));
    }
    // ---- DIISX reductions
    {
        FinBuilder fb;
        std::vector<std::vector<unsigned char>> sn_part(nh), sn_fin(nh), zd_part(nh), zd_fin(nh);
        for (int i = 0; i < nh; ++i) {
            rcc_t* h = p->hs[i];
            const int64_t n = amp_n(h);
            double2* d = h->diis_err[cpos[i]];
            sn_part[i].resize(batch_param_size(RK_SUMNORM_PART));
            sn_fin[i].resize(batch_param_size(RK_FINISH));
            sumnorm_params(d, n, h->red_dst, h->red_part, RED_BLOCKS, sn_part[i].data(),
                           sn_fin[i].data());
            std::vector<const double2*> xs(h->diis_err.begin(), h->diis_err.begin() + h->it);
            zd_part[i].resize(batch_param_size(RK_ZDOT_PART));
            zd_fin[i].resize(batch_param_size(RK_FINISH));
            zdot_params(xs.data(), h->it, d, n, h->red_dst + 2, h->red_part, RED_BLOCKS,
                        zd_part[i].data(), zd_fin[i].data());
        }
        fb.add(RK_DIIS_PREP, per_handle(p, RK_DIIS_PREP, [&](rcc_t* h, int i, void* q) {
                   diis_prep_params(amp_S(h), amp_T(h), h->diis_err[cpos[i]],
                                    h->diis_amp[cpos[i]], amp_n(h), q);
               }), nmax);
        fb.add(RK_SUMNORM_PART, sn_part, bmax);
        fb.add(RK_FINISH, sn_fin, 2);
        fb.add(RK_ZDOT_PART, zd_part, bmax);
        fb.add(RK_FINISH, zd_fin, itmax);
        fb.add(RK_COPY, per_handle(p, RK_COPY, [&](rcc_t* h, int i, void* q) {
                   copy_params(2 + h->it, h->red_dst, p->dslots + (size_t)i * RED_SLOTS, q);
               }), 2 + itmax);
        CK(fin_run(p, fb, slots));
        HIPCHK(h0, hipStreamSynchronize(p->st));
    }
    // ---- host FLIN, then extrapolation, CCENG on S and S -> T
    std::vector<std::vector<double>> coef(nh);
    for (int i = 0; i < nh; ++i) {
        rcc_t* h = p->hs[i];
        memcpy(h->hred, p->hslots + (size_t)i * RED_SLOTS, (2 + h->it) * sizeof(double2));
        CK(diis_coefs(h, rms ? rms + i : nullptr, coef[i]));
    }
    {
        FinBuilder fb;
        std::vector<std::vector<unsigned char>> e1p(nh), e1f(nh), e2p(nh), e2f(nh);
        int b1 = 1, b2 = 1;
        int64_t n2 = 0;
        for (int i = 0; i < nh; ++i) {
            rcc_t* h = p->hs[i];
            const double2* fvo = amp_FVO(h);
            const double2* vvoo = amp_VVOO(h);
            for (auto* v : {&e1p[i], &e2p[i]}) v->resize(batch_param_size(RK_ZDOT_PART));
            for (auto* v : {&e1f[i], &e2f[i]}) v->resize(batch_param_size(RK_FINISH));
            zdot_params(&fvo, 1, amp_S(h), amp_n1(h), h->red_dst + RED_E, h->red_part,
                        RED_BLOCKS, e1p[i].data(), e1f[i].data());
            zdot_params(&vvoo, 1, amp_TAU(h), amp_n2(h), h->red_dst + RED_E + 1, h->red_part,
                        RED_BLOCKS, e2p[i].data(), e2f[i].data());
            b1 = std::max(b1, red_blocks(amp_n1(h), RED_BLOCKS));
            b2 = std::max(b2, red_blocks(amp_n2(h), RED_BLOCKS));
            n2 = std::max<int64_t>(n2, amp_n2(h));
            h->tau_ok = false;     // TAU will hold tau(S), not tau(T)
        }
        fb.add(RK_LINCOMB, per_handle(p, RK_LINCOMB, [&](rcc_t* h, int i, void* q) {
                   std::vector<const double2*> amps(h->diis_amp.begin(),
                                                    h->diis_amp.begin() + h->it);
                   lincomb_params(amps.data(), coef[i].data(), h->it, amp_S(h), amp_n(h), q);
               }), nmax);
        // tau(S): the C1 kernel, or the blocked engine's bilinear gather
        const int tk = h0->bk ? RK_MAPBIL : RK_GETTAU;
        fb.add(tk, per_handle(p, tk, [&](rcc_t* h, int, void* q) {
                   if (h->bk) blk_tau_s_params(h->bk, q);
                   else gettau_params(h->S1(), h->S2(), h->TAU, 1.0, h->vv, h->oo, h->nv, q);
               }), n2);
        fb.add(RK_ZDOT_PART, e1p, b1);
        fb.add(RK_FINISH, e1f, 1);
        fb.add(RK_ZDOT_PART, e2p, b2);
        fb.add(RK_FINISH, e2f, 1);
        fb.add(RK_COPY, per_handle(p, RK_COPY, [&](rcc_t* h, int, void* q) {
                   copy_params(amp_n(h), amp_S(h), amp_T(h), q);
                   if (h->bk) h->c1t_stale = true;
               }), nmax);
        fb.add(RK_COPY, per_handle(p, RK_COPY, [&](rcc_t* h, int i, void* q) {
                   copy_params(2, h->red_dst + RED_E, p->dslots + (size_t)i * RED_SLOTS + RED_E, q);
               }), 2);
        CK(fin_run(p, fb, slots));
        HIPCHK(h0, hipEventRecord(p->et1, p->st));
        // the batch's work is complete: later work on the handles' own
        // streams needs no join
        HIPCHK(h0, hipStreamSynchronize(p->st));
    }
    float ms = 0.f;
    HIPCHK(h0, hipEventElapsedTime(&ms, p->et0, p->et1));
    for (int i = 0; i < nh; ++i) {
        rcc_t* h = p->hs[i];
        memcpy(h->hred + RED_E, p->hslots + (size_t)i * RED_SLOTS + RED_E, 2 * sizeof(double2));
        energy_end(h, ecc ? ecc + i : nullptr, nullptr);
        invalidate_t(h);
        h->last_iter_ms = ms;       // the batch iteration, the same for every handle
        h->nit += 1;
        if (nit) nit[i] = h->nit;
    }
    return 0;
}


// One grouped product over all handles: the batch stream waits for every
// handle's stream, runs the launch, and every handle's stream waits for it.
int batch_product(BatchPlan* p, int kind) {
    static const int opb[4] = {OP_N, OP_N, OP_C, OP_N};
    static const int tag[4] = {1, 0, 2, 2};
    rcc_t* h0 = p->hs[0];
    for (size_t i = 0; i < p->hs.size(); ++i) {
        HIPCHK(h0, hipEventRecord(p->evh[i], p->hs[i]->st));
        HIPCHK(h0, hipStreamWaitEvent(p->st, p->evh[i], 0));
    }
    if (zgemm_grouped(opb[kind], tag[kind], p->gauss, p->dG[kind], p->ng[kind], p->nb[kind],
                      p->st) != 0) {
        h0->err = "batch grouped GEMM launch failed";
        return -1;
    }
    HIPCHK(h0, hipEventRecord(p->evb, p->st));
    for (rcc_t* h : p->hs) HIPCHK(h, hipStreamWaitEvent(h->st, p->evb, 0));
    return 0;
}

// The term phase (HOINTM .. BINTM, DENOM) of every handle with the four
// grouped products; the same arithmetic as run_terms, X formed first.
int batch_terms_grouped(BatchPlan* p) {
    for (rcc_t* h : p->hs) {
        CK(prep(h));
        HIPCHK(h, hipEventRecord(h->ev0, h->st));
        CK(x_pre(h));
    }
    CK(batch_product(p, 1));
    for (rcc_t* h : p->hs) {
        CK(x_post(h));
        CK(do_hointm(h));
        CK(do_hvintm(h));
        CK(do_hovintm(h));
        CK(do_t1eqns(h));
        CK(do_gointm(h));
        CK(do_gvintm(h));
        CK(do_aintm(h));
        CK(hintm_pre(h, true));
    }
    CK(batch_product(p, 2));
    for (rcc_t* h : p->hs) {
        CK(hintm_post(h));
        CK(t2eqns_pre(h));
    }
    CK(batch_product(p, 3));
    for (rcc_t* h : p->hs) {
        CK(t2eqns_post(h));
        CK(bintm_pre(h));
    }
    CK(batch_product(p, 0));
    for (rcc_t* h : p->hs) {
        CK(bintm_post(h));
        k_denom(h->S1(), h->S2(), h->eps, h->any_frozen ? h->frozen : nullptr, h->vv, h->oo,
                h->nv, h->no, h->st);
    }
    return 0;
}

}  // namespace

// ======================================================================= C ABI
extern "C" {

int rcc_create(const rcc_sym_t* sym, rcc_t** out) {
    if (!out) return -1;
    *out = nullptr;
    rcc_t* h = new rcc_t();
    *out = h;
    if (!sym || sym->nrep < 1 || sym->nrep > 32 || !sym->no || !sym->nv ||
        (sym->nrep > 1 && !sym->multb)) {
        h->err = "rcc_create: need 1 <= nrep <= 32, no[], nv[] and (nrep > 1) multb";
        return -1;
    }
    h->no = 0;
    h->nv = 0;
    for (int i = 0; i < sym->nrep; ++i) {
        if (sym->no[i] < 0 || sym->nv[i] < 0) {
            h->err = "rcc_create: negative orbital count";
            return -1;
        }
        h->no += sym->no[i];
        h->nv += sym->nv[i];
    }
    if (h->no > 4096 || h->nv > 8192) {
        h->err = "rcc_create: unsupported orbital-space size";
        return -1;
    }
    h->noot = (int64_t)h->no * (h->no - 1) / 2;
    h->nvvt = (int64_t)h->nv * (h->nv - 1) / 2;
    h->nvo = (int64_t)h->no * h->nv;
    h->nt1 = h->nvo;
    h->nt2 = h->nvvt * h->noot;
    HIPCHK(h, hipGetDevice(&h->device));
    // address of this device's allocation-free scratch (rcc_zgemm_device_noalloc)
    if (zgemm_scratch_init(h->device) != 0) {
        h->err = "rcc_create: no address for the device's product scratch";
        return -1;
    }
    HIPCHK(h, hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
    HIPCHK(h, hipEventCreate(&h->ev0));
    HIPCHK(h, hipEventCreate(&h->ev1));
    for (int w = 0; w < 4; ++w)
        for (int p = 0; p < 2; ++p) HIPCHK(h, hipEventCreate(&h->kev[w][p]));
    CK(build_sym(h, sym));
    if (sym->nrep > 1) {
        // the blocked engine, unless RCC_BLOCKED=0 or the irreps are not their
        // own inverse (then the C1 working layout with grouped dominant products)
        const char* env = std::getenv("RCC_BLOCKED");
        if (!(env && env[0] == '0')) {
            blk::SymIn si;
            si.nrep = sym->nrep;
            si.no = h->sym.no;
            si.nv = h->sym.nv;
            si.ptab = h->sym.ptab;
            si.frozen.assign(h->no, 0);
            if (sym->frozen)
                for (int i = 0; i < h->no; ++i) si.frozen[i] = sym->frozen[i] ? 1 : 0;
            const int rc = blk_create(si, &h->bk, h->err);
            if (rc < 0) return -1;
            if (rc > 0) {
                h->bk = nullptr;
                h->err.clear();
            }
        }
    }
    CK(alloc_all(h));
    if (sym->nrep > 1) {
        CK(alloc_blocks(h));
        if (h->bk) {
            if (blk_bind(h->bk, blk::A_VVVV, h->VVVV_s, h->err) != 0 ||
                blk_bind(h->bk, blk::A_VOVV, h->VOVV_s, h->err) != 0)
                return -1;
        } else {
            CK(setup_grouped(h));
        }
        // off-symmetry blocks of the C1 arrays are never written
        k_fill_zero(h->FOO, (int64_t)h->no * h->no, h->st);
        k_fill_zero(h->FVO, h->nvo, h->st);
        k_fill_zero(h->FVV, (int64_t)h->nv * h->nv, h->st);
        HIPCHK(h, hipStreamSynchronize(h->st));
    }
    // pair tables
    std::vector<int2> oo, vv;
    for (int y = 0; y < h->no; ++y)
        for (int x = y + 1; x < h->no; ++x) oo.push_back(make_int2(x, y));
    for (int y = 0; y < h->nv; ++y)
        for (int x = y + 1; x < h->nv; ++x) vv.push_back(make_int2(x, y));
    CK(dalloc_raw(h, &h->oo_tab, (int64_t)std::max<size_t>(oo.size(), 1) * sizeof(int2)));
    CK(dalloc_raw(h, &h->vv_tab, (int64_t)std::max<size_t>(vv.size(), 1) * sizeof(int2)));
    if (!oo.empty()) HIPCHK(h, hipMemcpy(h->oo_tab, oo.data(), oo.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (!vv.empty()) HIPCHK(h, hipMemcpy(h->vv_tab, vv.data(), vv.size() * sizeof(int2), hipMemcpyHostToDevice));
    h->oo = PairTab{h->oo_tab, h->no, h->noot};
    h->vv = PairTab{h->vv_tab, h->nv, h->nvvt};
    std::vector<int> fr(std::max(h->no, 1), 0);
    if (sym->frozen)
        for (int i = 0; i < h->no; ++i) {
            fr[i] = sym->frozen[i] ? 1 : 0;
            h->any_frozen = h->any_frozen || fr[i];
        }
    HIPCHK(h, hipMemcpy(h->frozen, fr.data(), fr.size() * sizeof(int), hipMemcpyHostToDevice));
    CK(alloc_diis(h));
    return 0;
}

void rcc_destroy(rcc_t* h) {
    if (!h) return;
    drop_plans_of(h);
    if (h->st) (void)hipStreamSynchronize(h->st);
    drop_graph(h);
    if (h->t3) t3_free(h->t3);
    if (h->bk) blk_destroy(h->bk);
    if (h->hred) (void)hipHostFree(h->hred);
    for (void* p : h->allocs) (void)hipFree(p);
    for (int w = 0; w < 4; ++w)
        for (int p = 0; p < 2; ++p)
            if (h->kev[w][p]) (void)hipEventDestroy(h->kev[w][p]);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    for (hipEvent_t e : h->mev)
        if (e) (void)hipEventDestroy(e);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
}

const char* rcc_last_error(const rcc_t* h) { return h ? h->err.c_str() : "null handle"; }

int rcc_set_options(rcc_t* h, int maxdim, int maxit, int ntol, int precision) {
    if (!h) return -1;
    if (maxdim < 1 || maxdim > MAXDIM_CAP) {
        h->err = "rcc_set_options: maxdim must be 1.." + std::to_string(MAXDIM_CAP);
        return -1;
    }
    if (precision != 0 && h->sym.nrep > 1) {
        h->err = "rcc_set_options: split precision is available for C1 handles (nrep = 1)";
        return -1;
    }
    if (precision < 0 || precision > 2) {
        h->err = "rcc_set_options: precision must be 0 (FP64), 1 (split-BF16) or 2 (split-INT8)";
        return -1;
    }
    if (precision != h->precision) {
        drop_graph(h);
        free_split(h);
        invalidate_t(h);
    }
    h->maxdim = maxdim;
    h->maxit = maxit;
    h->ntol = ntol;
    h->precision = precision;
    return alloc_diis(h);
}

int rcc_set_split_digits(rcc_t* h, int digits) {
    if (!h) return -1;
    if (digits < 4 || digits > 5) {
        h->err = "rcc_set_split_digits: digits must be 4 or 5";
        return -1;
    }
    if (digits != h->digits) {
        drop_graph(h);
        h->digits = digits;
        h->split_const_ok = h->split_tau_ok = h->split_ring_ok = false;
        h->x_ok = false;
    }
    return 0;
}

int rcc_set_complex_product(rcc_t* h, int three_mult) {
    if (!h) return -1;
    const int g = three_mult ? 1 : 0;
    if (g != h->gemm_gauss) {
        // a captured iteration and any compiled batch plan hold the old
        // product's kernels: rebuild them on the next call
        drop_graph(h);
        drop_plans_of(h);
        h->gemm_gauss = g;
    }
    return 0;
}

int rcc_set_fock(rcc_t* h, const double* foo, const double* fvo, const double* fvv,
                 const double* eps) {
    if (!h) return -1;
    if (h->sym.nrep > 1) {
        CK(sym_to_c1(h, RCC_LAY_FOO, foo, hipMemcpyHostToDevice, h->FOO, (int64_t)h->no * h->no));
        CK(sym_to_c1(h, RCC_LAY_T1, fvo, hipMemcpyHostToDevice, h->FVO, h->nvo));
        CK(sym_to_c1(h, RCC_LAY_FVV, fvv, hipMemcpyHostToDevice, h->FVV, (int64_t)h->nv * h->nv));
        HIPCHK(h, hipMemcpy(h->eps, eps, (size_t)(h->no + h->nv) * 8, hipMemcpyHostToDevice));
        h->derived_ok = false;
        h->bk_const_ok = false;
        return 0;
    }
    CK(upload(h, h->FOO, foo, (int64_t)h->no * h->no, hipMemcpyHostToDevice));
    CK(upload(h, h->FVO, fvo, h->nvo, hipMemcpyHostToDevice));
    CK(upload(h, h->FVV, fvv, (int64_t)h->nv * h->nv, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->eps, eps, (size_t)(h->no + h->nv) * 8, hipMemcpyHostToDevice));
    return 0;
}

static int set_ints_common(rcc_t* h, int cls, int irrep, int64_t col0, int64_t ncols,
                           const void* buf, hipMemcpyKind kind) {
    if (!h) return -1;
    if (cls < 0 || cls > 5) {
        h->err = "rcc_set_ints: bad class";
        return -1;
    }
    if (irrep < 0 || irrep >= h->sym.nrep) {
        h->err = "rcc_set_ints: irrep out of range";
        return -1;
    }
    if (h->sym.nrep == 1) {
        if (col0 < 0 || ncols < 0 || col0 + ncols > h->cls_cols[cls]) {
            h->err = "rcc_set_ints: column batch outside the class";
            return -1;
        }
        const int64_t rows = h->cls_rows[cls];
        CK(upload(h, h->cls[cls] + col0 * rows, buf, rows * ncols, kind));
    } else {
        int rk, ck;
        int64_t ld;
        lay_spec(h, cls, &rk, &ck, &ld);
        const int64_t rows = lay_block_rows(h, rk, irrep), cols = lay_block_rows(h, ck, irrep);
        if (col0 < 0 || ncols < 0 || col0 + ncols > cols) {
            h->err = "rcc_set_ints: column batch outside the irrep block";
            return -1;
        }
        const double2* dsrc = (const double2*)buf;
        if (kind == hipMemcpyHostToDevice) {
            CK(ensure_stage(h, rows * ncols));
            HIPCHK(h, hipMemcpyAsync(h->stage, buf, (size_t)(rows * ncols) * 16, kind, h->st));
            dsrc = h->stage;
        }
        const SymInfo& S = h->sym;
        // the block copies: VVVV and VOVV always, the other classes in the
        // blocked engine (whose layouts are these)
        double2* blk = cls == RCC_VVVV ? h->VVVV_s : cls == RCC_VOVV ? h->VOVV_s
                       : h->bk ? blk_ptr(h->bk, blk::A_OOOO + cls) : nullptr;
        if (blk)
            HIPCHK(h, hipMemcpyAsync(blk + lay_offset(h, rk, ck, irrep) + col0 * rows, dsrc,
                                     (size_t)(rows * ncols) * 16, hipMemcpyDeviceToDevice, h->st));
        if (cls != RCC_VVVV && h->cls[cls])
            k_scatter_block(dsrc, rows, ncols, S.dmap + S.doff[rk][irrep],
                            S.dmap + S.doff[ck][irrep] + col0, h->cls[cls], ld, h->st);
        HIPCHK(h, hipStreamSynchronize(h->st));
    }
    h->derived_ok = false;
    h->bk_const_ok = false;
    h->split_const_ok = false;
    if (h->graph_ok || h->graph_warm) drop_graph(h);
    return 0;
}

int rcc_set_ints(rcc_t* h, int cls, int irrep, int64_t col0, int64_t ncols, const double* buf) {
    return set_ints_common(h, cls, irrep, col0, ncols, buf, hipMemcpyHostToDevice);
}

int rcc_set_ints_device(rcc_t* h, int cls, int irrep, int64_t col0, int64_t ncols,
                        const void* dbuf) {
    return set_ints_common(h, cls, irrep, col0, ncols, dbuf, hipMemcpyDeviceToDevice);
}

int rcc_set_amplitudes(rcc_t* h, const double* t1, const double* t2) {
    if (!h) return -1;
    if (h->bk) {
        // the blocked layout is the symmetry layout of the interface
        if (t1) HIPCHK(h, hipMemcpy(blk_ptr(h->bk, blk::A_T1), t1, amp_n1(h) * 16,
                                    hipMemcpyHostToDevice));
        if (t2) HIPCHK(h, hipMemcpy(blk_ptr(h->bk, blk::A_T2), t2, amp_n2(h) * 16,
                                    hipMemcpyHostToDevice));
        h->c1t_stale = true;
    } else if (h->sym.nrep > 1) {
        if (t1) CK(sym_to_c1(h, RCC_LAY_T1, t1, hipMemcpyHostToDevice, h->T1(), h->nt1));
        if (t2) CK(sym_to_c1(h, RCC_LAY_T2, t2, hipMemcpyHostToDevice, h->T2(), h->nt2));
    } else {
        if (t1) CK(upload(h, h->T1(), t1, h->nt1, hipMemcpyHostToDevice));
        if (t2) CK(upload(h, h->T2(), t2, h->nt2, hipMemcpyHostToDevice));
    }
    invalidate_t(h);
    // a new start, as CCEQN's entry (NIT = ITC = IT = 0, cceqns.F:255-260)
    h->nit = 0;
    h->it = 0;
    h->itc = 0;
    std::fill(h->bb2.begin(), h->bb2.end(), 0.0);
    return 0;
}

int rcc_get_amplitudes(rcc_t* h, double* t1, double* t2) {
    if (!h) return -1;
    HIPCHK(h, hipStreamSynchronize(h->st));
    if (h->bk) {
        if (t1) HIPCHK(h, hipMemcpy(t1, blk_ptr(h->bk, blk::A_T1), amp_n1(h) * 16,
                                    hipMemcpyDeviceToHost));
        if (t2) HIPCHK(h, hipMemcpy(t2, blk_ptr(h->bk, blk::A_T2), amp_n2(h) * 16,
                                    hipMemcpyDeviceToHost));
        return 0;
    }
    if (h->sym.nrep > 1) {
        if (t1) CK(c1_to_sym_host(h, RCC_LAY_T1, h->T1(), t1));
        if (t2) CK(c1_to_sym_host(h, RCC_LAY_T2, h->T2(), t2));
        return 0;
    }
    if (t1) HIPCHK(h, hipMemcpy(t1, h->T1(), h->nt1 * 16, hipMemcpyDeviceToHost));
    if (t2) HIPCHK(h, hipMemcpy(t2, h->T2(), h->nt2 * 16, hipMemcpyDeviceToHost));
    return 0;
}

int rcc_get_array(rcc_t* h, int which, double* out, int64_t n) {
    if (!h) return -1;
    if (h->bk) {
        int id = -1;
        switch (which) {
            case RCC_ARR_HO: id = blk::A_HO; break;
            case RCC_ARR_HV: id = blk::A_HV; break;
            case RCC_ARR_HOV: id = blk::A_HOV; break;
            case RCC_ARR_GO: id = blk::A_GO; break;
            case RCC_ARR_GV: id = blk::A_GV; break;
            case RCC_ARR_AZ: id = blk::A_AZ; break;
            case RCC_ARR_H: id = blk::A_H; break;
            case RCC_ARR_S1: id = blk::A_S1; break;
            case RCC_ARR_S2: id = blk::A_S2; break;
            case RCC_ARR_T1: id = blk::A_T1; break;
            case RCC_ARR_T2: id = blk::A_T2; break;
            case RCC_ARR_TAU: id = blk::A_TAU; break;
            case RCC_ARR_OOOO: id = blk::A_OOOO; break;
            case RCC_ARR_VOOO: id = blk::A_VOOO; break;
            case RCC_ARR_VVOO: id = blk::A_VVOO; break;
            case RCC_ARR_VOVO: id = blk::A_VOVO; break;
            case RCC_ARR_VOVV: id = blk::A_VOVV; break;
            case RCC_ARR_VVVV: id = blk::A_VVVV; break;
            default: break;
        }
        if (id >= 0) {
            const int64_t need = blk_size(h->bk, id);
            if (n != need) {
                h->err = "rcc_get_array: size mismatch (expected " + std::to_string(need) + ")";
                return -1;
            }
            HIPCHK(h, hipStreamSynchronize(h->st));
            HIPCHK(h, hipMemcpy(out, blk_ptr(h->bk, id), (size_t)need * 16,
                                hipMemcpyDeviceToHost));
            return 0;
        }
    }
    const double2* src = nullptr;
    int64_t len = 0;
    const int o = h->no, v = h->nv;
    switch (which) {
        case RCC_ARR_HO: src = h->HO; len = (int64_t)o * o; break;
        case RCC_ARR_HV: src = h->HV; len = (int64_t)v * v; break;
        case RCC_ARR_HOV: src = h->HOV; len = h->nvo; break;
        case RCC_ARR_GO: src = h->GO; len = (int64_t)o * o; break;
        case RCC_ARR_GV: src = h->GV; len = (int64_t)v * v; break;
        case RCC_ARR_AZ: src = h->AZ; len = h->noot * h->noot; break;
        case RCC_ARR_H: src = h->H; len = h->nvo * h->nvo; break;
        case RCC_ARR_S1: src = h->S1(); len = h->nt1; break;
        case RCC_ARR_S2: src = h->S2(); len = h->nt2; break;
        case RCC_ARR_T1: src = h->T1(); len = h->nt1; break;
        case RCC_ARR_T2: src = h->T2(); len = h->nt2; break;
        case RCC_ARR_TAU: src = h->TAU; len = h->nt2; break;
        case RCC_ARR_FOO: src = h->FOO; len = (int64_t)o * o; break;
        case RCC_ARR_FVO: src = h->FVO; len = h->nvo; break;
        case RCC_ARR_FVV: src = h->FVV; len = (int64_t)v * v; break;
        case RCC_ARR_OOOO: case RCC_ARR_VOOO: case RCC_ARR_VVOO: case RCC_ARR_VOVO:
        case RCC_ARR_VOVV: case RCC_ARR_VVVV: {
            const int c = which - RCC_ARR_OOOO;
            src = h->cls[c];
            len = h->cls_rows[c] * h->cls_cols[c];
            break;
        }
        case RCC_ARR_DOO: case RCC_ARR_DVV: case RCC_ARR_XVO: case RCC_ARR_DVO:
        case RCC_ARR_DOV: case RCC_ARR_MP2T1: case RCC_ARR_MP2T2: {
            if (!h->mp2d_ok) {
                h->err = "rcc_get_array: no MP2 density (rcc_mp2_density first)";
                return -1;
            }
            const double2* d = h->mp2d;
            const int64_t oo = (int64_t)o * o, vv = (int64_t)v * v, nvo = h->nvo;
            switch (which) {
                case RCC_ARR_DOO: src = d; len = oo; break;
                case RCC_ARR_DVV: src = d + oo; len = vv; break;
                case RCC_ARR_XVO: src = d + oo + vv; len = nvo; break;
                case RCC_ARR_DVO: src = d + oo + vv + nvo; len = nvo; break;
                case RCC_ARR_DOV: src = d + oo + vv + 2 * nvo; len = nvo; break;
                case RCC_ARR_MP2T1: src = h->mp2t; len = h->nt1; break;
                default: src = h->mp2t + h->nt1; len = h->nt2; break;
            }
            if (h->sym.nrep > 1) {
                const int lay = which == RCC_ARR_DOO   ? RCC_LAY_FOO
                                : which == RCC_ARR_DVV ? RCC_LAY_FVV
                                : which == RCC_ARR_MP2T2 ? RCC_LAY_T2 : RCC_LAY_T1;
                const int64_t need = lay_total(h, lay);
                if (n != need) {
                    h->err = "rcc_get_array: size mismatch (expected " + std::to_string(need) + ")";
                    return -1;
                }
                HIPCHK(h, hipStreamSynchronize(h->st));
                return c1_to_sym_host(h, lay, src, out);
            }
            if (n != len) {
                h->err = "rcc_get_array: size mismatch (expected " + std::to_string(len) + ")";
                return -1;
            }
            HIPCHK(h, hipStreamSynchronize(h->st));
            HIPCHK(h, hipMemcpy(out, src, (size_t)len * 16, hipMemcpyDeviceToHost));
            return 0;
        }
        default: h->err = "rcc_get_array: unknown array"; return -1;
    }
    if (which >= RCC_ARR_FOO && h->sym.nrep > 1) {
        // symmetry-blocked layout; VOVV and VVVV from their block copies
        const int lay = which == RCC_ARR_FOO ? RCC_LAY_FOO
                        : which == RCC_ARR_FVO ? RCC_LAY_T1
                        : which == RCC_ARR_FVV ? RCC_LAY_FVV : which - RCC_ARR_OOOO;
        const int64_t need = lay_total(h, lay);
        if (n != need) {
            h->err = "rcc_get_array: size mismatch (expected " + std::to_string(need) + ")";
            return -1;
        }
        HIPCHK(h, hipStreamSynchronize(h->st));
        if (which == RCC_ARR_VOVV || which == RCC_ARR_VVVV) {
            const double2* blk = which == RCC_ARR_VVVV ? h->VVVV_s : h->VOVV_s;
            HIPCHK(h, hipMemcpy(out, blk, (size_t)need * 16, hipMemcpyDeviceToHost));
            return 0;
        }
        return c1_to_sym_host(h, lay, src, out);
    }
    if (h->sym.nrep > 1) {
        int lay = RCC_LAY_T2;
        switch (which) {
            case RCC_ARR_HO: case RCC_ARR_GO: lay = RCC_LAY_FOO; break;
            case RCC_ARR_HV: case RCC_ARR_GV: lay = RCC_LAY_FVV; break;
            case RCC_ARR_HOV: case RCC_ARR_S1: case RCC_ARR_T1: lay = RCC_LAY_T1; break;
            case RCC_ARR_AZ: lay = RCC_OOOO; break;
            case RCC_ARR_H: lay = RCC_VOVO; break;
            default: break;
        }
        int rk, ck;
        int64_t ld;
        lay_spec(h, lay, &rk, &ck, &ld);
        len = lay_offset(h, rk, ck, h->sym.nrep);
        if (n != len) {
            h->err = "rcc_get_array: size mismatch (expected " + std::to_string(len) + ")";
            return -1;
        }
        HIPCHK(h, hipStreamSynchronize(h->st));
        return c1_to_sym_host(h, lay, src, out);
    }
    if (n != len) {
        h->err = "rcc_get_array: size mismatch (expected " + std::to_string(len) + ")";
        return -1;
    }
    HIPCHK(h, hipStreamSynchronize(h->st));
    HIPCHK(h, hipMemcpy(out, src, len * 16, hipMemcpyDeviceToHost));
    return 0;
}

int rcc_layout_block(rcc_t* h, int lay, int block, int64_t* rows, int64_t* cols,
                     int64_t* offset) {
    if (!h) return -1;
    int rk, ck;
    int64_t ld;
    if (!lay_spec(h, lay, &rk, &ck, &ld) || block < 0 || block > h->sym.nrep) {
        h->err = "rcc_layout_block: unknown layout or block";
        return -1;
    }
    if (rows) *rows = block < h->sym.nrep ? lay_block_rows(h, rk, block) : 0;
    if (cols) *cols = block < h->sym.nrep ? lay_block_rows(h, ck, block) : 0;
    if (offset) *offset = lay_offset(h, rk, ck, block);
    return 0;
}

int rcc_mp2(rcc_t* h, double* emp2) {
    if (!h) return -1;
    if (h->bk) {
        CK(bprep(h));
        CK(brun(h, blk::P_MP2));
        h->c1t_stale = true;
        invalidate_t(h);
        std::vector<double2> a, b;
        CK(zdot_multi(h, {amp_FVO(h)}, blk_ptr(h->bk, blk::A_T1), amp_n1(h), a));
        CK(zdot_multi(h, {amp_VVOO(h)}, blk_ptr(h->bk, blk::A_T2), amp_n2(h), b));
        if (emp2) *emp2 = a[0].x + b[0].x;
        h->nit = 0;
        h->it = 0;
        h->itc = 0;
        std::fill(h->bb2.begin(), h->bb2.end(), 0.0);
        return 0;
    }
    CK(prep(h));
    HIPCHK(h, hipMemcpyAsync(h->T1(), h->FVO, h->nt1 * 16, hipMemcpyDeviceToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(h->T2(), h->cls[RCC_VVOO], h->nt2 * 16, hipMemcpyDeviceToDevice, h->st));
    k_denom(h->T1(), h->T2(), h->eps, h->any_frozen ? h->frozen : nullptr, h->vv, h->oo, h->nv,
            h->no, h->st);
    invalidate_t(h);
    std::vector<double2> a, b;
    CK(zdot_multi(h, {h->FVO}, h->T1(), h->nt1, a));
    CK(zdot_multi(h, {h->cls[RCC_VVOO]}, h->T2(), h->nt2, b));
    if (emp2) *emp2 = a[0].x + b[0].x;
    h->nit = 0;
    h->it = 0;
    h->itc = 0;
    std::fill(h->bb2.begin(), h->bb2.end(), 0.0);
    return 0;
}

// one term: the blocked engine's program or the C1 routine
#define TERM(name, prog, call)                                  \
    int name(rcc_t* h) {                                        \
        if (!h) return -1;                                      \
        if (h->bk) {                                            \
            CK(bprep(h));                                       \
            return brun(h, prog);                               \
        }                                                       \
        return call;                                            \
    }
TERM(rcc_hointm, blk::P_HOINTM, do_hointm(h))
TERM(rcc_hvintm, blk::P_HVINTM, do_hvintm(h))
TERM(rcc_hovintm, blk::P_HOVINTM, do_hovintm(h))
TERM(rcc_t1eqns, blk::P_T1EQNS, do_t1eqns(h))
TERM(rcc_gointm, blk::P_GOINTM, do_gointm(h))
TERM(rcc_gvintm, blk::P_GVINTM, do_gvintm(h))
TERM(rcc_aintm, blk::P_AINTM, do_aintm(h))
TERM(rcc_t2eqns, blk::P_T2EQNS, do_t2eqns(h))
TERM(rcc_bintm, blk::P_BINTM, do_bintm(h))
#undef TERM

// HINTM as relccsd calls it: HAKCI formed here, not taken from GVINTM
int rcc_hintm(rcc_t* h) {
    if (!h) return -1;
    if (h->bk) {
        CK(bprep(h));
        CK(brun(h, blk::P_HINTM_HAKCI));
        return brun(h, blk::P_HINTM);
    }
    return do_hintm(h, false);
}

int rcc_denom(rcc_t* h) {
    if (!h) return -1;
    if (h->bk) {
        CK(bprep(h));
        return brun(h, blk::P_DIVIDE);
    }
    k_denom(h->S1(), h->S2(), h->eps, nullptr, h->vv, h->oo, h->nv, h->no, h->st);
    return 0;
}

int rcc_zcore(rcc_t* h) {
    if (!h) return -1;
    if (h->bk) {
        CK(bprep(h));
        return brun(h, blk::P_ZCORE);
    }
    if (h->any_frozen)
        k_denom(h->S1(), h->S2(), nullptr, h->frozen, h->vv, h->oo, h->nv, h->no, h->st);
    return 0;
}

int rcc_diis(rcc_t* h, double* rms) { return h ? do_diis(h, rms) : -1; }

int rcc_energy(rcc_t* h, double* ecc) { return h ? do_energy(h, ecc, nullptr) : -1; }

int rcc_t1diag(rcc_t* h, double* diag) {
    return h ? do_t1diag(h, amp_S(h), amp_n1(h), diag) : -1;
}

int rcc_accept(rcc_t* h) {
    if (!h) return -1;
    HIPCHK(h, hipMemcpyAsync(amp_T(h), amp_S(h), amp_n(h) * 16, hipMemcpyDeviceToDevice, h->st));
    if (h->bk) h->c1t_stale = true;
    invalidate_t(h);
    return 0;
}

int rcc_iterate(rcc_t* h, int* nit, double* ecc, double* rms) {
    if (!h) return -1;
    double e = 0, r = 0;
    CK(do_iteration(h, &e, &r));
    if (nit) *nit = h->nit;
    if (ecc) *ecc = e;
    if (rms) *rms = r;
    return 0;
}

int rcc_set_graph(rcc_t* h, int enable) {
    if (!h) return -1;
    drop_graph(h);
    h->use_graph = enable != 0;
    return 0;
}

int rcc_iterate_batch(rcc_t* const* hs, int n, int* nit, double* ecc, double* rms) {
    if (!hs || n < 0) return -1;
    for (int i = 0; i < n; ++i) {
        if (!hs[i]) return -1;
        HIPCHK(hs[i], hipEventRecord(hs[i]->ev0, hs[i]->st));
        CK(terms_phase(hs[i]));      // all systems' term phases in flight together
    }
    // the host waits once per stage for all systems
    for (int i = 0; i < n; ++i) CK(finish_a(hs[i]));
    for (int i = 0; i < n; ++i) HIPCHK(hs[i], hipStreamSynchronize(hs[i]->st));
    for (int i = 0; i < n; ++i) CK(finish_b(hs[i], rms ? rms + i : nullptr));
    for (int i = 0; i < n; ++i) HIPCHK(hs[i], hipStreamSynchronize(hs[i]->st));
    for (int i = 0; i < n; ++i) {
        CK(finish_c(hs[i], ecc ? ecc + i : nullptr));
        if (nit) nit[i] = hs[i]->nit;
    }
    return 0;
}

int rcc_iterate_batch_grouped(rcc_t* const* hs, int n, int* nit, double* ecc, double* rms) {
    if (!hs || n < 0) return -1;
    for (int i = 0; i < n; ++i)
        if (!hs[i]) return -1;
    if (n == 0) return 0;
    if (!batch_groupable(hs, n)) {
        hs[0]->err = "rcc_iterate_batch_grouped: every handle must have precision 0, the same "
                     "complex-product form and, with symmetry, self-inverse irreps";
        return -1;
    }
    BatchPlan* p = nullptr;
    CK(get_plan(hs, n, &p));
    bool graph = true;
    for (int i = 0; i < n; ++i) graph = graph && hs[i]->use_graph;
    if (p->prog) {
        int rc = batch_program_phase(p, graph);
        if (rc == 0) rc = batch_finish(p, nit, ecc, rms);
        if (rc != 0) (void)hipStreamSynchronize(p->st);   // no batch work left in flight
        return rc;
    }
    bool any_bk = false;
    for (int i = 0; i < n; ++i) any_bk = any_bk || hs[i]->bk;
    if (any_bk) {
        // blocked handles (alone or among others) without a compiled program:
        // every handle's own term phase
        for (int i = 0; i < n; ++i) {
            HIPCHK(hs[i], hipEventRecord(hs[i]->ev0, hs[i]->st));
            CK(terms_phase(hs[i]));
        }
    } else {
        CK(batch_terms_grouped(p));
    }
    for (int i = 0; i < n; ++i) CK(finish_a(hs[i]));
    for (int i = 0; i < n; ++i) HIPCHK(hs[i], hipStreamSynchronize(hs[i]->st));
    for (int i = 0; i < n; ++i) CK(finish_b(hs[i], rms ? rms + i : nullptr));
    for (int i = 0; i < n; ++i) HIPCHK(hs[i], hipStreamSynchronize(hs[i]->st));
    for (int i = 0; i < n; ++i) {
        CK(finish_c(hs[i], ecc ? ecc + i : nullptr));
        if (nit) nit[i] = hs[i]->nit;
    }
    return 0;
}

int rcc_batch_plan_info(rcc_t* const* hs, int n, int* info) {
    if (!hs || n <= 0 || !info) return -1;
    for (int i = 0; i < 8; ++i) info[i] = 0;
    std::vector<rcc_t*> key(hs, hs + n);
    for (BatchPlan* p : g_plans)
        if (p->hs == key) {
            info[0] = p->prog ? 1 : 0;
            if (p->prog) batch_info(p->prog, info + 1, info + 2, info + 3);
            return 0;
        }
    hs[0]->err = "rcc_batch_plan_info: no plan for these handles (run rcc_iterate_batch_grouped)";
    return -1;
}

int rcc_cceqn(rcc_t* h, int* nit, double* ecc, double* rms) {
    if (!h) return -1;
    const double conv = std::pow(10.0, -h->ntol);
    double e = 0, r = 1.0;
    int n = 0;
    while (n < h->maxit) {
        CK(do_iteration(h, &e, &r));
        ++n;
        if (r < conv) break;
    }
    if (nit) *nit = n;
    if (ecc) *ecc = e;
    if (rms) *rms = r;
    return 0;
}

int rcc_set_autotune(rcc_t* h, int enable) {
    if (!h) return -1;
    zgemm_autotune(enable);
    return 0;
}

int rcc_autotune_entries(rcc_t* h, int64_t* fp64_shapes, int64_t* int8_shapes) {
    if (!h) return -1;
    if (fp64_shapes) *fp64_shapes = zgemm_tuned_shapes();
    if (int8_shapes) *int8_shapes = i8_tuned_shapes();
    return 0;
}

// ---- integral front end: MRCONEE + MDCINT -> classes and Fock (RDINTS) ----

// Spinor classification of SYMTAB (ccdriv.F:832-858): within each
// higher-group irrep, in file order, the first NELEC spinors are occupied.
// C1 working position: occupied irrep-major then virtual irrep-major.
static int molfdir_positions(rcc_t* h, const MolfdirOne& one, const int* nelec, int nnelec,
                             std::vector<int>& spos, std::vector<int>& occ_sp,
                             std::vector<int>& vir_sp) {
    if (nnelec != one.nsymrp) {
        h->err = "rcc_read_molfdir: " + std::to_string(nnelec) + " NELEC values for " +
                 std::to_string(one.nsymrp) + " irreps in MRCONEE";
        return -1;
    }
    const SymInfo& S = h->sym;
    if (one.nrep != S.nrep) {
        h->err = "rcc_read_molfdir: MRCONEE has " + std::to_string(one.nrep) +
                 " abelian irreps, the handle " + std::to_string(S.nrep);
        return -1;
    }
    const int nrep = S.nrep;
    std::vector<int> seen(one.nsymrp + 1, 0), no_r(nrep, 0), nv_r(nrep, 0);
    std::vector<char> occ(one.nsp, 0);
    for (int s = 0; s < one.nsp; ++s) {
        const int g = one.irrep_higher[s], a = one.irrep[s];
        if (g < 1 || g > one.nsymrp || a < 1 || a > nrep) {
            h->err = "rcc_read_molfdir: spinor irrep out of range in MRCONEE";
            return -1;
        }
        occ[s] = ++seen[g] <= nelec[g - 1];
        if (occ[s]) ++no_r[a - 1];
        else ++nv_r[a - 1];
    }
    for (int r = 0; r < nrep; ++r)
        if (no_r[r] != S.no[r] || nv_r[r] != S.nv[r]) {
            h->err = "rcc_read_molfdir: irrep " + std::to_string(r + 1) + " has " +
                     std::to_string(no_r[r]) + " occupied / " + std::to_string(nv_r[r]) +
                     " virtual spinors in the files, the handle " + std::to_string(S.no[r]) +
                     " / " + std::to_string(S.nv[r]);
            return -1;
        }
    // C1 working position: occupied irrep-major then virtual irrep-major, file
    // order inside an irrep (SYMTAB's IREPSPI, ccdriv.F:832-858)
    std::vector<int> fo(nrep, 0), fv(nrep, 0);
    spos.assign(one.nsp, 0);
    occ_sp.assign(h->no, 0);
    vir_sp.assign(h->nv, 0);
    for (int s = 0; s < one.nsp; ++s) {
        const int r = one.irrep[s] - 1;
        if (occ[s]) {
            const int p = S.io[r] + fo[r]++;
            spos[s] = p;
            occ_sp[p] = s;
        } else {
            const int p = S.iv[r] + fv[r]++;
            spos[s] = -p - 1;
            vir_sp[p] = s;
        }
    }
    return 0;
}

// breit (may be null): the MDBINT list, sorted with ITRSGN = -1 into zeroed
// classes, to which the Coulomb classes are then added (RDINTS ccints.F:
// 230-238, 354-358, 457-461: the Breit pass ends with XAXPY(+1, Coulomb)).
static int sort_upload_and_run(rcc_t* h, const std::vector<int>& spos, const int* kr, int nkr,
                               const int4* d_idx, const double2* d_val, int64_t m, double* ms,
                               const MdcintList* breit = nullptr) {
    const int nsp = (int)spos.size();
    if (2 * nkr != nsp) {
        h->err = "rcc_sort: INCONSISTENT MRCONEE AND MDCINT FILES (2 NKR != NSP)";
        return -1;
    }
    std::vector<int> krt(2 * (size_t)nkr + 1, 0);
    for (int i = 1; i <= nkr; ++i) {
        const int a = kr[2 * (i - 1)], b = kr[2 * (i - 1) + 1];
        if (a < 1 || a > nsp || b < 1 || b > nsp) {
            h->err = "rcc_sort: KR table entry outside 1..NSP";
            return -1;
        }
        krt[nkr + i] = a - 1;
        krt[nkr - i] = b - 1;
    }
    CK(ensure_c1(h));
    int *d_spos = nullptr, *d_kr = nullptr;
    HIPCHK(h, hipMalloc((void**)&d_spos, spos.size() * 4));
    HIPCHK(h, hipMalloc((void**)&d_kr, krt.size() * 4));
    HIPCHK(h, hipMemcpy(d_spos, spos.data(), spos.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(d_kr, krt.data(), krt.size() * 4, hipMemcpyHostToDevice));
    // symmetric handles keep VVVV only in its irrep blocks: sort it into a
    // C1-sized scratch array and gather the blocks from there
    double2* cls[6];
    for (int c = 0; c < 6; ++c) cls[c] = h->cls[c];
    double2* vvvv_c1 = nullptr;
    if (h->sym.nrep > 1)
        HIPCHK(h, hipMalloc((void**)&vvvv_c1, (size_t)h->nvvt * h->nvvt * sizeof(double2) + 16));
    if (vvvv_c1) cls[RCC_VVVV] = vvvv_c1;
    int rc = sort_integrals_device(h->no, h->nv, cls, d_spos, nsp, d_kr, nkr, d_idx, d_val,
                                   m, h->st, ms, h->err);
    if (!rc && breit) {
        int64_t sizes[6], total = 0;
        for (int c = 0; c < 6; ++c) {
            sizes[c] = h->cls_rows[c] * h->cls_cols[c];
            total += sizes[c];
        }
        const int64_t mb = (int64_t)breit->idx.size();
        double2 *saved = nullptr, *b_val = nullptr;
        int4* b_idx = nullptr;
        hipError_t e = hipMalloc((void**)&saved, (size_t)std::max<int64_t>(total, 1) * 16);
        if (e == hipSuccess) e = hipMalloc((void**)&b_idx, (size_t)std::max<int64_t>(mb, 1) * 16);
        if (e == hipSuccess) e = hipMalloc((void**)&b_val, (size_t)std::max<int64_t>(mb, 1) * 16);
        if (e == hipSuccess && mb) {
            e = hipMemcpy(b_idx, breit->idx.data(), (size_t)mb * 16, hipMemcpyHostToDevice);
            if (e == hipSuccess)
                e = hipMemcpy(b_val, breit->val.data(), (size_t)mb * 16, hipMemcpyHostToDevice);
        }
        if (e != hipSuccess) {
            h->err = std::string("Breit list: ") + hipGetErrorString(e);
            rc = -1;
        } else {
            int64_t off = 0;
            for (int c = 0; c < 6; ++c) {
                k_copy(sizes[c], cls[c], saved + off, h->st);
                off += sizes[c];
            }
            double ms_b = 0.0;
            rc = sort_integrals_device(h->no, h->nv, cls, d_spos, nsp, d_kr, nkr, b_idx, b_val, mb,
                                       h->st, &ms_b, h->err, -1);
            if (!rc) {
                off = 0;
                for (int c = 0; c < 6; ++c) {
                    k_axpby(sizes[c], cplx(1, 0), saved + off, cplx(1, 0), cls[c], h->st);
                    off += sizes[c];
                }
                if (hipStreamSynchronize(h->st) != hipSuccess) rc = -1;
                if (ms) *ms += ms_b;
            }
        }
        (void)hipFree(saved);
        (void)hipFree(b_idx);
        (void)hipFree(b_val);
    }
    if (!rc && h->sym.nrep > 1) {
        c1_to_sym_dev(h, RCC_VOVV, h->cls[RCC_VOVV], h->VOVV_s);
        c1_to_sym_dev(h, RCC_VVVV, vvvv_c1, h->VVVV_s);
        if (h->bk)
            for (int c : {RCC_OOOO, RCC_VOOO, RCC_VVOO, RCC_VOVO})
                c1_to_sym_dev(h, c, h->cls[c], blk_ptr(h->bk, blk::A_OOOO + c));
        (void)hipStreamSynchronize(h->st);
    }
    (void)hipFree(d_spos);
    (void)hipFree(d_kr);
    if (vvvv_c1) (void)hipFree(vvvv_c1);
    if (rc) return -1;
    h->derived_ok = false;
    h->bk_const_ok = false;
    h->split_const_ok = false;
    h->tau_ok = h->x_ok = h->ring_ok = false;
    if (h->graph_ok || h->graph_warm) drop_graph(h);
    return 0;
}

int rcc_sort_ints_device(rcc_t* h, const int* spinor_pos, int nsp, const int* kr, int nkr,
                         const void* d_idx, const void* d_val, int64_t m, double* ms) {
    if (!h || !spinor_pos || !kr) return -1;
    if (h->sym.nrep != 1) {
        h->err = "rcc_sort_ints_device: C1 handles only";
        return -1;
    }
    std::vector<int> spos(spinor_pos, spinor_pos + nsp);
    int no = 0, nv = 0;
    for (int s = 0; s < nsp; ++s) {
        if (spos[s] >= 0) ++no;
        else ++nv;
    }
    std::vector<char> seen_o(h->no, 0), seen_v(h->nv, 0);
    bool ok = no == h->no && nv == h->nv;
    for (int s = 0; ok && s < nsp; ++s) {
        const int p = spos[s] >= 0 ? spos[s] : -spos[s] - 1;
        std::vector<char>& seen = spos[s] >= 0 ? seen_o : seen_v;
        ok = p < (int)seen.size() && !seen[p];
        if (ok) seen[p] = 1;
    }
    if (!ok) {
        h->err = "rcc_sort_ints_device: spinor positions are not one occupied position 0..no-1 "
                 "and one virtual position 0..nv-1 each";
        return -1;
    }
    return sort_upload_and_run(h, spos, kr, nkr, (const int4*)d_idx, (const double2*)d_val, m, ms);
}

int rcc_read_molfdir(rcc_t* h, const char* mrconee, const char* mdcint, const int* nelec,
                     int nnelec, double* e0e1) {
    return rcc_read_molfdir_opts(h, mrconee, mdcint, nelec, nnelec, 0, 0, e0e1);
}

int rcc_read_molfdir_opts(rcc_t* h, const char* mrconee, const char* mdcint, const int* nelec,
                          int nnelec, int useoe, int norecmp, double* e0e1) {
    if (!h || !mrconee || !mdcint || !nelec) return -1;
    MolfdirOne one;
    if (read_mrconee(mrconee, one, h->err)) return -1;
    std::vector<int> spos, occ_sp, vir_sp;
    CK(molfdir_positions(h, one, nelec, nnelec, spos, occ_sp, vir_sp));
    MdcintList lst;
    if (read_mdcint(mdcint, lst, h->err)) return -1;
    const int64_t m = (int64_t)lst.idx.size();
    int4* d_idx = nullptr;
    double2* d_val = nullptr;
    HIPCHK(h, hipMalloc((void**)&d_idx, (size_t)std::max<int64_t>(m, 1) * sizeof(int4)));
    HIPCHK(h, hipMalloc((void**)&d_val, (size_t)std::max<int64_t>(m, 1) * sizeof(double2)));
    if (m) {
        HIPCHK(h, hipMemcpy(d_idx, lst.idx.data(), (size_t)m * sizeof(int4), hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(d_val, lst.val.data(), (size_t)m * sizeof(double2), hipMemcpyHostToDevice));
    }
    // BREIT in MRCONEE: the Breit (Gaunt) list MDBINT beside MDCINT is sorted
    // too and added (RDINTS ccints.F:496-505)
    MdcintList blst;
    if (one.breit) {
        std::string bp(mdcint);
        const size_t slash = bp.find_last_of('/');
        bp = (slash == std::string::npos ? std::string() : bp.substr(0, slash + 1)) + "MDBINT";
        if (read_mdcint(bp.c_str(), blst, h->err, true)) {
            (void)hipFree(d_idx);
            (void)hipFree(d_val);
            return -1;
        }
        if (blst.nkr != lst.nkr) {
            (void)hipFree(d_idx);
            (void)hipFree(d_val);
            h->err = "MDCINT AND MDBINT FILES ARE INCOMPATIBLE (NKR differs)";
            return -1;
        }
    }
    double ms = 0.0;
    const int rc = sort_upload_and_run(h, spos, lst.kr.data(), lst.nkr, d_idx, d_val, m, &ms,
                                       one.breit ? &blst : nullptr);
    (void)hipFree(d_idx);
    (void)hipFree(d_val);
    if (rc) return -1;
    // Fock blocks and orbital energies (GETFM / MAKEFM; USEOE = F)
    double2* d_h1 = nullptr;
    int *d_occ = nullptr, *d_vir = nullptr;
    double* d_e01 = nullptr;
    const size_t nh = (size_t)one.nsp * one.nsp;
    HIPCHK(h, hipMalloc((void**)&d_h1, nh * sizeof(double2)));
    HIPCHK(h, hipMalloc((void**)&d_occ, occ_sp.size() * 4 + 4));
    HIPCHK(h, hipMalloc((void**)&d_vir, vir_sp.size() * 4 + 4));
    HIPCHK(h, hipMalloc((void**)&d_e01, 2 * sizeof(double)));
    HIPCHK(h, hipMemcpy(d_h1, one.h1.data(), nh * sizeof(double2), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(d_occ, occ_sp.data(), occ_sp.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(d_vir, vir_sp.data(), vir_sp.size() * 4, hipMemcpyHostToDevice));
    const int fr = build_fock_device(h->no, h->nv, one.nsp, d_h1, d_occ, d_vir, h->cls, h->FOO,
                                     h->FVO, h->FVV, h->eps, d_e01, h->st, h->err);
    double e01[2] = {0.0, 0.0};
    if (!fr) HIPCHK(h, hipMemcpyAsync(e01, d_e01, 2 * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    (void)hipFree(d_h1);
    (void)hipFree(d_occ);
    (void)hipFree(d_vir);
    (void)hipFree(d_e01);
    if (fr) return -1;
    // namelist CCSORT (GETFM ccints.F:1080, 1150-1198, 1260, 1287): USEOE keeps
    // MRCONEE's orbital energies; NORECMP takes F = diag of them, FVO = 0 and
    // E0 = their occupied sum (E1 comes from the integrals either way)
    if (useoe || norecmp) {
        const int o = h->no, v = h->nv;
        std::vector<double> ew((size_t)o + v);
        for (int p = 0; p < o; ++p) ew[p] = one.eps[occ_sp[p]];
        for (int p = 0; p < v; ++p) ew[o + p] = one.eps[vir_sp[p]];
        HIPCHK(h, hipMemcpy(h->eps, ew.data(), ew.size() * 8, hipMemcpyHostToDevice));
        if (norecmp) {
            std::vector<double2> foo((size_t)o * o, make_double2(0.0, 0.0)),
                fvo((size_t)v * o, make_double2(0.0, 0.0)), fvv((size_t)v * v, make_double2(0.0, 0.0));
            double e0 = 0.0;
            for (int p = 0; p < o; ++p) {
                foo[p + (size_t)o * p].x = ew[p];
                e0 += ew[p];
            }
            for (int p = 0; p < v; ++p) fvv[p + (size_t)v * p].x = ew[o + p];
            HIPCHK(h, hipMemcpy(h->FOO, foo.data(), foo.size() * 16, hipMemcpyHostToDevice));
            HIPCHK(h, hipMemcpy(h->FVO, fvo.data(), fvo.size() * 16, hipMemcpyHostToDevice));
            HIPCHK(h, hipMemcpy(h->FVV, fvv.data(), fvv.size() * 16, hipMemcpyHostToDevice));
            e01[0] = e0;
        }
    }
    h->derived_ok = false;
    h->bk_const_ok = false;
    h->last_sort_ms = ms;
    if (e0e1) {
        e0e1[0] = e01[0];
        e0e1[1] = e01[1];
        e0e1[2] = one.ecore;
    }
    return 0;
}

double rcc_last_sort_ms(rcc_t* h) { return h ? h->last_sort_ms : -1.0; }
int rcc_read_property(rcc_t* h, const char* mrconee, const char* mdprop, const char* name,
                      const int* nelec, int nnelec, double* prop) {
    if (!h || !mrconee || !mdprop || !name || !nelec || !prop) return -1;
    MolfdirOne one;
    if (read_mrconee(mrconee, one, h->err)) return -1;
    std::vector<int> spos, occ_sp, vir_sp;
    CK(molfdir_positions(h, one, nelec, nnelec, spos, occ_sp, vir_sp));
    std::vector<double2> p;
    if (read_mdprop(mdprop, name, one.nsp, p, h->err)) return -1;
    // the operator in the handle's spinor order: occupied positions, then virtual
    const int n = h->no + h->nv;
    std::vector<int> sp(n);
    for (int i = 0; i < h->no; ++i) sp[i] = occ_sp[i];
    for (int a = 0; a < h->nv; ++a) sp[h->no + a] = vir_sp[a];
    double2* out = (double2*)prop;
    for (int q = 0; q < n; ++q)
        for (int r = 0; r < n; ++r) out[r + (int64_t)q * n] = p[sp[r] + (int64_t)sp[q] * one.nsp];
    return 0;
}

int rcc_expectation(rcc_t* h, const double* prop, int level, double* value) {
    if (!h || !prop || !value) return -1;
    if (level < 0 || level > 2) {
        h->err = "rcc_expectation: level 0 (reference determinant), 1 (CCSD, unrelaxed) or 2 "
                 "(MP2, relaxed)";
        return -1;
    }
    if (level == 2 && !h->mp2d_ok) {
        h->err = "rcc_expectation: level 2 needs rcc_mp2_density first";
        return -1;
    }
    const int n = h->no + h->nv;
    const double2* P = (const double2*)prop;
    double re = 0.0, im = 0.0;
    for (int i = 0; i < h->no; ++i) {
        re += P[i + (int64_t)i * n].x;
        im += P[i + (int64_t)i * n].y;
    }
    if (level == 1) {
        std::vector<double2> D((size_t)n * n);
        CK(rcc_density(h, (double*)D.data()));
        // tr(D P) = sum_pq D(p,q) P(q,p)
        for (int q = 0; q < n; ++q)
            for (int p = 0; p < n; ++p) {
                const double2 d = D[p + (int64_t)q * n], a = P[q + (int64_t)p * n];
                re += d.x * a.x - d.y * a.y;
                im += d.x * a.y + d.y * a.x;
            }
    }
    if (level == 2) {
        // EXPVAL (ccgrad.F:300-333): unconjugated dot products of the operator
        // blocks EOV(A,I) = P(I,A), EVO(A,I) = P(A,I), EOO, EVV with DVO, DOV,
        // DOO, DVV
        const int o = h->no, v = h->nv;
        const int64_t nvo = h->nvo;
        std::vector<double2> d((size_t)o * o + (size_t)v * v + 3 * nvo);
        HIPCHK(h, hipStreamSynchronize(h->st));
        HIPCHK(h, hipMemcpy(d.data(), h->mp2d, d.size() * 16, hipMemcpyDeviceToHost));
        const double2* doo = d.data();
        const double2* dvv = doo + (size_t)o * o;
        const double2* dvo = dvv + (size_t)v * v + nvo;
        const double2* dov = dvo + nvo;
        auto acc = [&](double2 p, double2 q) {
            re += p.x * q.x - p.y * q.y;
            im += p.x * q.y + p.y * q.x;
        };
        for (int i = 0; i < o; ++i)
            for (int a = 0; a < v; ++a) {
                const size_t ai = a + (size_t)v * i;
                acc(P[i + (int64_t)(o + a) * n], dvo[ai]);
                acc(P[(o + a) + (int64_t)i * n], dov[ai]);
            }
        for (int j = 0; j < o; ++j)
            for (int i = 0; i < o; ++i) acc(P[i + (int64_t)j * n], doo[i + (size_t)o * j]);
        for (int b = 0; b < v; ++b)
            for (int a = 0; a < v; ++a)
                acc(P[(o + a) + (int64_t)(o + b) * n], dvv[a + (size_t)v * b]);
    }
    value[0] = re;
    value[1] = im;
    return 0;
}


// ---- MP2 relaxed density: the DOMP2G block of CCFOPR (ccdriv.F:2058-2090) ----
namespace {

int free_tracked(rcc_t* h, void* p) {
    if (!p) return 0;
    auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
    if (it != h->allocs.end()) h->allocs.erase(it);
    HIPCHK(h, hipFree(p));
    return 0;
}

// RPAA's DIISX step (cceqns.F:1985-2178) on the joint vector (X, Y) with
// Y = conj X, which every iterate keeps (PRECD divides by a real DG and the
// extrapolation coefficients are real): with d the error of X, the joint
// sums are sum d + conj sum d = 2 Re sum d and 2 sum |d|^2, and the joint
// B entries are Re + Im of <d_k|d> + conj <d_k|d> = 2 Re <d_k|d>.
struct ZDiis {
    int md = 8, it = 0, itc = 0;
    std::vector<double> bb2;
};

int zdiis_step(rcc_t* h, ZDiis& z, double2** err, double2** amp, const double2* cur, double2* nw,
               int64_t n, double* rms) {
    hipStream_t st = h->st;
    z.it = std::min(z.it + 1, z.md);
    z.itc = z.itc + 1 > z.md ? 1 : z.itc + 1;
    const int c = z.itc - 1;
    double2* d = err[c];
    k_copy(n, nw, d, st);
    k_axpby(n, cplx(-1, 0), cur, cplx(1, 0), d, st);
    k_copy(n, nw, amp[c], st);
    k_sum_and_norm(d, n, h->red_dst, h->red_part, RED_BLOCKS, st);
    std::vector<const double2*> xs(err, err + z.it);
    k_zdotc_multi(xs.data(), z.it, d, n, h->red_dst + 2, h->red_part, RED_BLOCKS, st);
    HIPCHK(h, hipMemcpyAsync(h->hred, h->red_dst, (2 + z.it) * sizeof(double2),
                             hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipStreamSynchronize(st));
    const double2 sum = h->hred[0];
    const double sum2 = 2.0 * h->hred[1].x, csum = 2.0 * sum.x;
    *rms = std::sqrt(std::fabs(sum2 - csum * csum));
    const int md = z.md, ld = md + 1;
    auto BB2 = [&](int i, int j) -> double& { return z.bb2[i + (int64_t)j * ld]; };
    for (int k = 0; k < z.it; ++k) {
        const double vv = 2.0 * h->hred[2 + k].x;
        BB2(k, c) = vv;
        BB2(c, k) = vv;
    }
    const int it = z.it;
    std::vector<double> bb((size_t)ld * (md + 2), 0.0);
    auto BB = [&](int i, int j) -> double& { return bb[i + (int64_t)j * ld]; };
    double xm = 0.0;
    for (int i = 0; i < it; ++i)
        for (int j = 0; j < it; ++j) {
            BB(i, j) = BB2(i, j);
            xm = std::max(xm, std::fabs(BB(i, j)));
        }
    // every stored error is zero (e.g. a zero right-hand side): the bordered
    // system is singular, and the plain step is already the fixed point
    if (!(xm > 0.0)) return 0;
    for (int i = 0; i < it; ++i)
        for (int j = 0; j < it; ++j) BB(i, j) /= xm;
    for (int i = 0; i < it; ++i) {
        BB(i, it) = -1.0;
        BB(it, i) = -1.0;
    }
    BB(it, it) = 0.0;
    BB(it, it + 1) = -1.0;
    flin(bb.data(), ld, it + 1, 1);
    std::vector<double> coef(it);
    for (int i = 0; i < it; ++i) coef[i] = BB(i, it + 1);
    std::vector<const double2*> as(amp, amp + it);
    k_lincomb(as.data(), coef.data(), it, nw, n, st);
    return 0;
}

}  // namespace

int rcc_mp2_density(rcc_t* h, int maxit, int maxdim, int ntol, int* nit, double* rms) {
    if (h) CK(ensure_c1(h));
    if (!h) return -1;
    h->mp2d_ok = false;
    // Symmetry-blocked handles run in the C1 working layout, where every
    // class but VVVV is held with its symmetry-forbidden elements zero: the
    // densities and X come out zero outside their totally symmetric blocks,
    // and the Z-vector's right-hand side, hence its solution, too.
    if (maxit < 1 || maxdim < 1 || maxdim > 16 || ntol < 1 || ntol > 16) {
        h->err = "rcc_mp2_density: maxit >= 1, 1 <= maxdim <= 16, 1 <= ntol <= 16";
        return -1;
    }
    CK(prep(h));
    const int o = h->no, v = h->nv;
    const int64_t nvo = h->nvo, noot = h->noot, nvvt = h->nvvt, nt1 = h->nt1, nt2 = h->nt2;
    hipStream_t st = h->st;
    if (!h->mp2t) {
        CK(dalloc(h, &h->mp2t, nt1 + nt2));
        CK(dalloc(h, &h->mp2d, (int64_t)o * o + (int64_t)v * v + 3 * nvo));
        CK(dalloc_raw(h, &h->zdg, std::max<int64_t>(nvo, 1) * 8));
    }
    const int nch = zvec_chunks(nvo);
    // the Z-vector work, and before it the partial sums of the X contractions
    const int64_t need = std::max({(2 * (int64_t)maxdim + nch + 1) * std::max<int64_t>(nvo, 1),
                                   2 * noot * v, 2 * nvvt * o, (int64_t)1});
    if (h->zwork_n < need) {
        CK(free_tracked(h, h->zwork));
        h->zwork = nullptr;
        h->zwork_n = 0;
        CK(dalloc(h, &h->zwork, need));
        h->zwork_n = need;
    }
    for (int k = 0; k < 8; ++k)
        if (!h->mev[k]) HIPCHK(h, hipEventCreate(&h->mev[k]));
    HIPCHK(h, hipEventRecord(h->mev[0], st));
    double2* T1m = h->mp2t;
    double2* T2m = h->mp2t + nt1;
    double2* DOO = h->mp2d;
    double2* DVV = DOO + (int64_t)o * o;
    double2* XVO = DVV + (int64_t)v * v;
    double2* DVO = XVO + nvo;
    double2* DOV = DVO + nvo;
    // MP2EQN 'LAMBDA' (cceqns.F:89-92, 116-117): T = (FVO, VVOO) / D and the
    // lambda amplitudes equal to T -- L below is T itself
    k_copy(nt1, h->FVO, T1m, st);
    k_copy(nt2, h->cls[RCC_VVOO], T2m, st);
    k_denom(T1m, T2m, h->eps, h->any_frozen ? h->frozen : nullptr, h->vv, h->oo, v, o, st);
    // CCDENSC(1) (ccgrad.F:77-104)
    // DOO = conj[ -T(CDL,I)^H L(CDL,J) - T(C,I)^H L(C,J) ]
    k_expand_e2(T2m, h->E2, nvvt, h->oo, o, st);
    {
        const int64_t K = nvvt * o;
        CK(gemm(h, OP_C, OP_N, o, o, (int)K, cplx(-1, 0), h->E2, K, h->E2, K, cplx(0, 0), DOO, o));
        CK(gemm(h, OP_C, OP_N, o, o, v, cplx(-1, 0), T1m, v, T1m, v, cplx(1, 0), DOO, o));
        k_conj((int64_t)o * o, DOO, DOO, st);
    }
    // DVV = conj[ T(A,DKL) L(B,DKL)^H + T(A,K) L(B,K)^H ]
    k_expand_f1(T2m, h->F1, noot, h->vv, v, st);
    {
        const int64_t K = (int64_t)v * noot;
        CK(gemm(h, OP_N, OP_C, v, v, (int)K, cplx(1, 0), h->F1, v, h->F1, v, cplx(0, 0), DVV, v));
        CK(gemm(h, OP_N, OP_C, v, v, o, cplx(1, 0), T1m, v, T1m, v, cplx(1, 0), DVV, v));
        k_conj((int64_t)v * v, DVV, DVV, st);
    }
    HIPCHK(h, hipEventRecord(h->mev[1], st));
    // CCDENSZ's X intermediate (ccgrad.F:199-229)
    k_fill_zero(XVO, nvo, st);
    // T1EQN1 with tau -> L: X(A,I) += V(K,L,C,I) L(A,C,K,L)   (F1 holds L expanded)
    CK(gemm(h, OP_N, OP_N, v, o, (int)(v * noot), cplx(1, 0), h->F1, v, h->M1, v * noot,
            cplx(1, 0), XVO, v));
    // T1EQN2 with tau -> L: X(A,I) += V(A,K,C,D) L(C,D,I,K)
    HIPCHK(h, hipEventRecord(h->mev[6], st));
    CK(gemm(h, OP_N, OP_N, (int)nvo, (int)noot, (int)nvvt, cplx(1, 0), h->cls[RCC_VOVV], nvo, T2m,
            nvvt, cplx(0, 0), h->X, nvo));
    HIPCHK(h, hipEventRecord(h->mev[7], st));
    h->x_ok = false;     // h->X no longer holds VOVV tau
    k_trace_t1eqn2(h->X, XVO, h->oo, v, o, st);
    // X(A,I) += V(IK,AL) D(K,L)   (SRT1ST4 + XGEMV)
    k_x_vooo_doo(h->cls[RCC_VOOO], DOO, XVO, h->oo, h->zwork, v, o, st);
    // X(A,I) += W(CI,DA) D(C,D)   (SRT20D, left-hand contraction)
    k_x_vovv_dvv(h->cls[RCC_VOVV], DVV, XVO, h->vv, h->zwork, v, o, st);
    k_axpby(nvo, cplx(0, 0), XVO, cplx(-1, 0), XVO, st);
    HIPCHK(h, hipEventRecord(h->mev[2], st));
    // RPAA at zero frequency (ccrpa.F:104-229): X1 = XVO / DG, then
    // X2 = (XVO + A' X1 - B conj X1) / DG, DIIS, until RMS < 10^-ntol
    k_zvec_dg(h->VOVO_R, h->eps, h->zdg, v, o, st);
    k_zvec_precd(XVO, h->zdg, DVO, nvo, st);
    std::vector<double2*> zerr(maxdim), zamp(maxdim);
    for (int k = 0; k < maxdim; ++k) {
        zerr[k] = h->zwork + (int64_t)k * nvo;
        zamp[k] = h->zwork + (int64_t)(maxdim + k) * nvo;
    }
    double2* part = h->zwork + 2 * (int64_t)maxdim * nvo;
    double2* X2 = part + (int64_t)nch * nvo;
    ZDiis z;
    z.md = maxdim;
    z.bb2.assign((size_t)(maxdim + 1) * (maxdim + 2), 0.0);
    const double tol = std::pow(10.0, -ntol);
    double r = 1.0, gemv_ms = 0.0;
    int it = 0;
    while (nvo > 0 && it < maxit) {
        ++it;
        HIPCHK(h, hipEventRecord(h->mev[4], st));
        k_zvec_gemv(h->VOVO_R, h->VVOO_R, DVO, part, nvo, st);
        HIPCHK(h, hipEventRecord(h->mev[5], st));
        k_zvec_step(part, XVO, h->VOVO_R, DVO, h->zdg, X2, nvo, st);
        CK(zdiis_step(h, z, zerr.data(), zamp.data(), DVO, X2, nvo, &r));   // synchronises
        float gms = 0.f;
        HIPCHK(h, hipEventElapsedTime(&gms, h->mev[4], h->mev[5]));
        gemv_ms += gms;
        k_copy(nvo, X2, DVO, st);
        if (r < tol) break;
    }
    k_conj(nvo, DVO, DOV, st);
    HIPCHK(h, hipEventRecord(h->mev[3], st));
    HIPCHK(h, hipEventSynchronize(h->mev[3]));
    float ms[3] = {0.f, 0.f, 0.f}, tot = 0.f;
    HIPCHK(h, hipEventElapsedTime(&tot, h->mev[0], h->mev[3]));
    for (int k = 0; k < 3; ++k) HIPCHK(h, hipEventElapsedTime(&ms[k], h->mev[k], h->mev[k + 1]));
    h->mp2_ms[0] = tot;
    for (int k = 0; k < 3; ++k) h->mp2_ms[k + 1] = ms[k];
    h->mp2_ms[4] = gemv_ms;
    float xg = 0.f;
    HIPCHK(h, hipEventElapsedTime(&xg, h->mev[6], h->mev[7]));
    h->mp2_ms[5] = xg;
    h->last_mp2d_ms = tot;
    h->mp2_nit = it;
    h->mp2_rms = r;
    h->mp2d_ok = true;
    if (nit) *nit = it;
    if (rms) *rms = r;
    return 0;
}

double rcc_last_mp2d_ms(rcc_t* h) { return h ? h->last_mp2d_ms : 0.0; }

int rcc_mp2d_timing(rcc_t* h, double* ms) {
    if (!h || !ms) return -1;
    for (int k = 0; k < 6; ++k) ms[k] = h->mp2_ms[k];
    return 0;
}

int rcc_get_eps(rcc_t* h, double* eps) {
    if (!h || !eps) return -1;
    HIPCHK(h, hipStreamSynchronize(h->st));
    HIPCHK(h, hipMemcpy(eps, h->eps, (size_t)(h->no + h->nv) * 8, hipMemcpyDeviceToHost));
    return 0;
}

// ---- MCCRES restart file of the handle's amplitudes (PUTAMPT / GETAMPT) ----
extern "C" int rcc_mccres_write(const char* path, int64_t ndimt1, int64_t ndimt2, const double* t1,
                                const double* t2, const int* imstat, const double* energies);
extern "C" int rcc_mccres_read(const char* path, int64_t ndimt1, int64_t ndimt2, double* t1,
                               double* t2, int* imstat, double* energies);
extern "C" const char* rcc_file_error(void);

int rcc_write_mccres(rcc_t* h, const char* path, const int* imstat, const double* energies) {
    if (!h || !path) return -1;
    const int64_t n1 = h->sym.nrep > 1 ? lay_total(h, RCC_LAY_T1) : h->nt1;
    const int64_t n2 = h->sym.nrep > 1 ? lay_total(h, RCC_LAY_T2) : h->nt2;
    std::vector<double> t1(2 * (size_t)n1), t2(2 * (size_t)n2);
    CK(rcc_get_amplitudes(h, t1.data(), t2.data()));
    if (rcc_mccres_write(path, n1, n2, t1.data(), t2.data(), imstat, energies)) {
        h->err = rcc_file_error();
        return -1;
    }
    return 0;
}

int rcc_read_mccres(rcc_t* h, const char* path, int* imstat, double* energies) {
    if (!h || !path) return -1;
    const int64_t n1 = h->sym.nrep > 1 ? lay_total(h, RCC_LAY_T1) : h->nt1;
    const int64_t n2 = h->sym.nrep > 1 ? lay_total(h, RCC_LAY_T2) : h->nt2;
    std::vector<double> t1(2 * (size_t)n1), t2(2 * (size_t)n2);
    if (rcc_mccres_read(path, n1, n2, t1.data(), t2.data(), imstat, energies)) {
        h->err = rcc_file_error();
        return -1;
    }
    return rcc_set_amplitudes(h, t1.data(), t2.data());
}

int64_t rcc_t3_triples(rcc_t* h) { return h ? t3_triples(h->no) : -1; }

int rcc_t3corr_part(rcc_t* h, int64_t first, int64_t count, double* et) {
    if (!h || !et) return -1;
    CK(ensure_c1(h));
    for (int c : {RCC_VOOO, RCC_VVOO, RCC_VOVV})
        if (!h->cls[c]) {
            h->err = "rcc_t3corr: integral class " + std::to_string(c) + " not set";
            return -1;
        }
    if (!h->t3) h->t3 = t3_new();
    CK(c1_sync_t(h));
    T3Input in;
    in.no = h->no;
    in.nv = h->nv;
    in.gauss = h->gemm_gauss;
    in.T1 = h->T1();
    in.T2 = h->T2();
    in.FVO = h->FVO;
    in.VOOO = h->cls[RCC_VOOO];
    in.VVOO = h->cls[RCC_VVOO];
    in.VOVV = h->cls[RCC_VOVV];
    in.eps = h->eps;
    if (count < 0) count = t3_triples(h->no) - first;
    return t3_run(h->t3, in, first, count, h->st, et, &h->t3_tm, h->err);
}

int rcc_t3corr(rcc_t* h, double* et) { return rcc_t3corr_part(h, 0, -1, et); }

double rcc_last_t3_ms(rcc_t* h, double* parts) {
    if (!h) return -1.0;
    if (parts) {
        parts[0] = h->t3_tm.gemm_ms;
        parts[1] = h->t3_tm.energy_ms;
        parts[2] = (double)h->t3_tm.chunk;
        parts[3] = (double)h->t3_tm.chunks;
    }
    return h->t3_tm.total_ms;
}

int rcc_lambda(rcc_t* h, int* nit, double* rms) {
    if (!h) return -1;
    return lam_solve(h, nit, rms);
}

// Lambda in / out in the handle's layouts (symmetry-blocked for nrep > 1,
// like the amplitudes).
static int lam_upload(rcc_t* h, const double* l1, const double* l2, double2* dst) {
    if (h->sym.nrep > 1) {
        k_fill_zero(dst, h->nt1 + h->nt2, h->st);
        CK(sym_to_c1(h, RCC_LAY_T1, l1, hipMemcpyHostToDevice, dst, h->nt1));
        CK(sym_to_c1(h, RCC_LAY_T2, l2, hipMemcpyHostToDevice, dst + h->nt1, h->nt2));
        return 0;
    }
    CK(upload(h, dst, l1, h->nt1, hipMemcpyHostToDevice));
    return upload(h, dst + h->nt1, l2, h->nt2, hipMemcpyHostToDevice);
}

static int lam_download(rcc_t* h, const double2* src, double* l1, double* l2) {
    HIPCHK(h, hipStreamSynchronize(h->st));
    if (h->sym.nrep > 1) {
        CK(c1_to_sym_host(h, RCC_LAY_T1, src, l1));
        return c1_to_sym_host(h, RCC_LAY_T2, src + h->nt1, l2);
    }
    HIPCHK(h, hipMemcpy(l1, src, (size_t)h->nt1 * 16, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(l2, src + h->nt1, (size_t)h->nt2 * 16, hipMemcpyDeviceToHost));
    return 0;
}

int rcc_set_lambda(rcc_t* h, const double* l1, const double* l2) {
    if (!h || !l1 || !l2) return -1;
    CK(lam_check(h));
    CK(prep(h));
    CK(lam_alloc(h));
    CK(lam_upload(h, l1, l2, h->LAM));
    h->lam_set = true;
    return 0;
}

int rcc_get_lambda(rcc_t* h, double* l1, double* l2) {
    if (!h || !l1 || !l2) return -1;
    if (!h->lam_set) {
        h->err = "rcc_get_lambda: no Lambda (call rcc_lambda or rcc_set_lambda)";
        return -1;
    }
    return lam_download(h, h->LAM, l1, l2);
}

int rcc_lambda_vjp(rcc_t* h, const double* l1, const double* l2, int with_energy, double* g1,
                   double* g2) {
    if (!h || !l1 || !l2 || !g1 || !g2) return -1;
    CK(lam_setup(h));
    double2* in = h->LAMS;
    double2* out = h->S;        // scratch once the forward pass is done
    CK(lam_upload(h, l1, l2, in));
    CK(lam_sweep(h, in, out, with_energy != 0));
    return lam_download(h, out, g1, g2);
}

int rcc_density(rcc_t* h, double* dens) {
    if (!h || !dens) return -1;
    if (!h->lam_set) {
        h->err = "rcc_density: no Lambda (call rcc_lambda or rcc_set_lambda)";
        return -1;
    }
    return lam_density(h, dens);
}

int rcc_last_lambda_ms(rcc_t* h, double* ms) {
    if (!h || !ms) return -1;
    *ms = h->last_lambda_ms;
    return 0;
}

int rcc_last_iteration_ms(rcc_t* h, double* ms) {
    if (!h || !ms) return -1;
    *ms = h->last_iter_ms;
    return 0;
}

int rcc_kernel_times(rcc_t* h, double* ms4) {
    if (!h || !ms4) return -1;
    for (int w = 0; w < 4; ++w) ms4[w] = h->kms[w];
    return 0;
}

int rcc_zgemm_device(rcc_t* h, int opa, int opb, int m, int n, int k, const double* alpha,
                     const void* A, int64_t lda, const void* B, int64_t ldb, const double* beta,
                     void* C, int64_t ldc) {
    if (!h) return -1;
    const bool conj_pair = (opa == OP_R && opb == OP_N) || (opa == OP_N && opb == OP_R);
    if (opa < 0 || opb < 0 || ((opa > 2 || opb > 2) && !conj_pair)) {
        h->err = "rcc_zgemm_device: op must be 0, 1 or 2 (3 only with 0 on the other operand)";
        return -1;
    }
    CK(gemm(h, opa, opb, m, n, k, cplx(alpha[0], alpha[1]), (const double2*)A, lda,
            (const double2*)B, ldb, cplx(beta[0], beta[1]), (double2*)C, ldc));
    HIPCHK(h, hipStreamSynchronize(h->st));
    return 0;
}

int rcc_zgemm_device_noalloc(rcc_t* h, int opa, int opb, int m, int n, int k, const double* alpha,
                             const void* A, int64_t lda, const void* B, int64_t ldb,
                             const double* beta, void* C, int64_t ldc, int ksplit) {
    if (!h) return -1;
    const bool conj_pair = (opa == OP_R && opb == OP_N) || (opa == OP_N && opb == OP_R);
    if (opa < 0 || opb < 0 || ((opa > 2 || opb > 2) && !conj_pair)) {
        h->err = "rcc_zgemm_device_noalloc: op must be 0, 1 or 2 (3 only with 0 on the other "
                 "operand)";
        return -1;
    }
    if (ksplit < 0) {
        h->err = "rcc_zgemm_device_noalloc: ksplit must be >= 0";
        return -1;
    }
    GemmArgs g;
    g.gauss = h->gemm_gauss;
    g.opA = opa; g.opB = opb; g.m = m; g.n = n; g.k = k;
    g.A = (const double2*)A; g.lda = lda;
    g.B = (const double2*)B; g.ldb = ldb;
    g.C = (double2*)C; g.ldc = ldc;
    g.alpha = cplx(alpha[0], alpha[1]);
    g.beta = cplx(beta[0], beta[1]);
    std::string err;
    if (zgemm_noalloc(g, h->device, ksplit, h->st, err) != 0) {
        h->err = "rcc_zgemm_device_noalloc: " + err;
        return -1;
    }
    return 0;
}

int rcc_zgemm_split_device(rcc_t* h, int m, int n, int k, const double* alpha, const void* A,
                           int64_t lda, const void* B, int64_t ldb, const double* beta, void* C,
                           int64_t ldc) {
    if (!h) return -1;
    SplitOperand a, b;
    split_layout(a, m, k);
    split_layout(b, n, k);
    void* mem[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // room for an 8-way K-split's partial sums
    const int64_t ws_elems = 8 * (int64_t)std::max(m, 1) * std::max(n, 1);
    HIPCHK(h, hipMalloc(&mem[0], std::max<int64_t>(1, a.plane_bytes)));
    HIPCHK(h, hipMalloc(&mem[1], std::max<int64_t>(1, a.ex_bytes)));
    HIPCHK(h, hipMalloc(&mem[2], std::max<int64_t>(1, b.plane_bytes)));
    HIPCHK(h, hipMalloc(&mem[3], std::max<int64_t>(1, b.ex_bytes)));
    HIPCHK(h, hipMalloc(&mem[4], ws_elems * sizeof(double2)));
    a.planes = mem[0];
    a.ex = (signed char*)mem[1];
    b.planes = mem[2];
    b.ex = (signed char*)mem[3];
    int rc = split_operand(a, (const double2*)A, lda, false, false, h->st);
    // B is k x n column-major: as a split operand its rows are n, k contiguous
    if (rc == 0) rc = split_operand(b, (const double2*)B, ldb, true, false, h->st);
    if (rc == 0)
        rc = zgemm_split(a, b, (double2*)C, ldc, cplx(alpha[0], alpha[1]), cplx(beta[0], beta[1]),
                         (double2*)mem[4], ws_elems, h->st);
    if (rc != 0) h->err = "rcc_zgemm_split_device: launch failed";
    HIPCHK(h, hipStreamSynchronize(h->st));
    for (void* p : mem) HIPCHK(h, hipFree(p));
    return rc;
}

int rcc_zgemm_i8_device(rcc_t* h, int digits, int m, int n, int k, const double* alpha,
                        const void* A, int64_t lda, const void* B, int64_t ldb,
                        const double* beta, void* C, int64_t ldc) {
    if (!h) return -1;
    if (digits < 4 || digits > 5 || m < 0 || n < 0 || k <= 0) {
        h->err = "rcc_zgemm_i8_device: digits must be 4 or 5, m, n >= 0, k > 0";
        return -1;
    }
    I8Operand a, b;
    i8_layout(a, m, k, digits);
    i8_layout(b, n, k, digits);
    void* mem[4] = {nullptr, nullptr, nullptr, nullptr};
    HIPCHK(h, hipMalloc(&mem[0], std::max<int64_t>(16, a.plane_bytes)));
    HIPCHK(h, hipMalloc(&mem[1], std::max<int64_t>(16, a.mx_elems * 8)));
    HIPCHK(h, hipMalloc(&mem[2], std::max<int64_t>(16, b.plane_bytes)));
    HIPCHK(h, hipMalloc(&mem[3], std::max<int64_t>(16, b.mx_elems * 8)));
    a.planes = mem[0];
    a.mx = (unsigned long long*)mem[1];
    b.planes = mem[2];
    b.mx = (unsigned long long*)mem[3];
    int rc = i8_split(a, (const double2*)A, lda, false, false, h->st);
    // B is k x n column-major: as a split operand its rows are n, k contiguous
    if (rc == 0) rc = i8_split(b, (const double2*)B, ldb, true, false, h->st);
    if (rc == 0)
        rc = zgemm_i8(a, b, (double2*)C, ldc, cplx(alpha[0], alpha[1]), cplx(beta[0], beta[1]),
                      h->st);
    if (rc != 0) h->err = "rcc_zgemm_i8_device: launch failed";
    HIPCHK(h, hipStreamSynchronize(h->st));
    for (void* p : mem) HIPCHK(h, hipFree(p));
    return rc;
}

int rcc_probe_mfma(rcc_t* h, double* out256) {
    if (!h) return -1;
    double* d = nullptr;
    HIPCHK(h, hipMalloc((void**)&d, 256 * 8));
    if (mfma_probe(d, h->st) != 0) {
        h->err = "probe launch failed";
        (void)hipFree(d);
        return -1;
    }
    HIPCHK(h, hipStreamSynchronize(h->st));
    HIPCHK(h, hipMemcpy(out256, d, 256 * 8, hipMemcpyDeviceToHost));
    HIPCHK(h, hipFree(d));
    return 0;
}

int rcc_probe_mfma_rate(rcc_t* h, int chains, int iters, int nblocks, double* tflops) {
    if (!h || !tflops) return -1;
    double* d = nullptr;
    HIPCHK(h, hipMalloc((void**)&d, 256 * 8));
    HIPCHK(h, hipEventRecord(h->ev0, h->st));
    if (mfma_rate(chains, iters, nblocks, d, h->st) != 0) {
        h->err = "rcc_probe_mfma_rate: chains must be 1, 2, 4, 8 or 12";
        (void)hipFree(d);
        return -1;
    }
    HIPCHK(h, hipEventRecord(h->ev1, h->st));
    HIPCHK(h, hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    HIPCHK(h, hipFree(d));
    // 4 waves per block, each MFMA 16*16*4 FMAs = 2048 flops
    *tflops = 2048.0 * 4.0 * chains * (double)iters * nblocks / (ms * 1e-3) / 1e12;
    return 0;
}

int rcc_synchronize(rcc_t* h) {
    if (!h) return -1;
    HIPCHK(h, hipStreamSynchronize(h->st));
    return 0;
}

}  // extern "C"
