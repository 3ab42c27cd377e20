// Host-side plan of the symmetry-blocked CCSD iteration (no HIP code here).
//
// With an abelian double group whose irreps are all their own inverse (the
// Z2^k double groups of relccsd: D2h* and its subgroups), every quantity of
// the CCSD equations is totally symmetric, so an array indexed by a row space
// R and a column space S is nonzero only where the irrep of the row equals the
// irrep of the column.  Stored by irrep g as a |R_g| x |S_g| column-major block
// (rows and columns in their C1 order), every array holds only its nonzero
// elements, and a contraction C(R,S) = sum_K A(R,K) B(K,S) is block-diagonal:
// C_g = A_g B_g for every g, one grouped GEMM.  relccsd does the same with its
// per-irrep CNTRCT calls (ccintm.F:1662-1723) on SYMTAB's layouts
// (ccdriv.F:746-1425), which these blocked layouts reproduce for the classes,
// the amplitudes, the Fock blocks and the intermediates HO .. H.
//
// The CCSD iteration becomes a fixed program of four kinds of operations:
// grouped products (descriptor tables with byte offsets into the arrays),
// gathers (out = beta out + alpha sum_t +-conj?(src_t[map_t]), index maps built
// here from the C1 meaning of each re-ordering relccsd's SRT* sorts perform),
// bilinear gathers (tau and the ring intermediate: products of two T1
// elements) and the denominators.  The device executor (rcc_blocked.hip) runs
// the program; nothing here depends on the values.
#pragma once
#include <algorithm>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

namespace rcc {
namespace blk {

enum Atom { AT_O = 0, AT_V = 1, AT_OO = 2, AT_VV = 3 };

// Symmetry of one handle: nrep fermion irreps, pair-irrep table ptab[i + n j]
// (0 the totally symmetric boson irrep), spinor counts per irrep, frozen flags
// of the occupied spinors in C1 order (occupied irrep-major, then virtual).
struct SymIn {
    int nrep = 1;
    std::vector<int> no, nv, ptab, frozen;
};

// Group elements: bosons 0..n-1, fermions n..2n-1; product table from ptab.
struct Group {
    int n = 1;
    std::vector<int> mul;
    int prod(int a, int b) const { return mul[(size_t)a * 2 * n + b]; }
};

// Builds the product table of the double group from the pair table, for
// groups whose irreps are self-inverse (ptab(i, i) = 0): boson b = f0 f_phi(b),
// b1 b2 = ptab(phi(b1), phi(b2)), f_i b = f_phi(ptab(0, i) b).  False when the
// table is not of that kind or the result is not an abelian group.
inline bool make_group(const SymIn& s, Group& G) {
    const int n = s.nrep;
    G.n = n;
    G.mul.assign((size_t)4 * n * n, -1);
    auto pt = [&](int i, int j) { return s.ptab[(size_t)i + (size_t)n * j]; };
    for (int i = 0; i < n; ++i)
        if (pt(i, i) != 0) return false;
    std::vector<int> phi(n, -1);
    for (int j = 0; j < n; ++j) {
        const int b = pt(0, j);
        if (b < 0 || b >= n || phi[b] >= 0) return false;
        phi[b] = j;
    }
    const int N = 2 * n;
    for (int b1 = 0; b1 < n; ++b1)
        for (int b2 = 0; b2 < n; ++b2) G.mul[(size_t)b1 * N + b2] = pt(phi[b1], phi[b2]);
    for (int i = 0; i < n; ++i)
        for (int b = 0; b < n; ++b) {
            const int f = n + phi[G.mul[(size_t)pt(0, i) * N + b]];
            G.mul[(size_t)(n + i) * N + b] = f;
            G.mul[(size_t)b * N + n + i] = f;
        }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) G.mul[(size_t)(n + i) * N + n + j] = pt(i, j);
    for (int a = 0; a < N; ++a)
        for (int b = 0; b < N; ++b) {
            if (G.prod(a, b) != G.prod(b, a)) return false;
            for (int c = 0; c < N; ++c)
                if (G.prod(G.prod(a, b), c) != G.prod(a, G.prod(b, c))) return false;
        }
    for (int a = 0; a < N; ++a)
        if (G.prod(0, a) != a || G.prod(a, a) != 0) return false;
    return true;
}

inline int64_t pair_ix(int x, int y, int m) {   // x > y, y outer / x inner
    return (int64_t)y * (2 * m - y - 1) / 2 + (x - y - 1);
}

// A composite index space: atoms first-fastest (C1 index = sum atom * radix).
struct Space {
    std::vector<int> at;
    std::vector<int64_t> rad;
    int64_t n1 = 1;
    std::vector<uint8_t> el;               // group element of each C1 index
    std::vector<int32_t> pos;              // position within its element's list
    std::vector<int64_t> cnt;              // per element
    std::vector<std::vector<int32_t>> list;   // per element: C1 indices ascending
};

struct Ctx {
    SymIn s;
    Group G;
    int o = 0, v = 0;
    int64_t noot = 0, nvvt = 0;
    std::vector<int> eo, ev;               // group element of each occupied / virtual spinor
    std::vector<int> oox, ooy, vvx, vvy;   // pair decode
    std::vector<int> eoo, evv;             // pair elements
};

inline void make_ctx(const SymIn& s, const Group& G, Ctx& c) {
    c.s = s;
    c.G = G;
    const int n = s.nrep;
    c.o = c.v = 0;
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < s.no[i]; ++k) c.eo.push_back(n + i);
        for (int k = 0; k < s.nv[i]; ++k) c.ev.push_back(n + i);
        c.o += s.no[i];
        c.v += s.nv[i];
    }
    c.noot = (int64_t)c.o * (c.o - 1) / 2;
    c.nvvt = (int64_t)c.v * (c.v - 1) / 2;
    for (int y = 0; y < c.o; ++y)
        for (int x = y + 1; x < c.o; ++x) {
            c.oox.push_back(x);
            c.ooy.push_back(y);
            c.eoo.push_back(G.prod(c.eo[x], c.eo[y]));
        }
    for (int y = 0; y < c.v; ++y)
        for (int x = y + 1; x < c.v; ++x) {
            c.vvx.push_back(x);
            c.vvy.push_back(y);
            c.evv.push_back(G.prod(c.ev[x], c.ev[y]));
        }
}

inline int64_t atom_size(const Ctx& c, int a) {
    return a == AT_O ? c.o : a == AT_V ? c.v : a == AT_OO ? c.noot : c.nvvt;
}

inline int atom_el(const Ctx& c, int a, int64_t x) {
    return a == AT_O ? c.eo[x] : a == AT_V ? c.ev[x] : a == AT_OO ? c.eoo[x] : c.evv[x];
}

inline void make_space(const Ctx& c, const std::vector<int>& atoms, Space& S) {
    S.at = atoms;
    S.rad.clear();
    S.n1 = 1;
    for (int a : atoms) {
        S.rad.push_back(atom_size(c, a));
        S.n1 *= atom_size(c, a);
    }
    const int N = 2 * c.G.n;
    S.el.assign((size_t)S.n1, 0);
    S.pos.assign((size_t)S.n1, 0);
    S.cnt.assign(N, 0);
    S.list.assign(N, {});
    std::vector<int64_t> dig(atoms.size(), 0);
    for (int64_t x = 0; x < S.n1; ++x) {
        int e = 0;
        for (size_t t = 0; t < atoms.size(); ++t) e = c.G.prod(e, atom_el(c, atoms[t], dig[t]));
        S.el[x] = (uint8_t)e;
        S.pos[x] = (int32_t)S.cnt[e]++;
        S.list[e].push_back((int32_t)x);
        for (size_t t = 0; t < atoms.size(); ++t) {
            if (++dig[t] < S.rad[t]) break;
            dig[t] = 0;
        }
    }
}

// A blocked array: row space R, column space S, block g = |R_g| x |S_g|.
struct Arr {
    std::string name;
    int R = -1, S = -1;
    std::vector<int64_t> off;   // per element; off[2n] = total
    int64_t size = 0;
    bool real = false;          // doubles (denominators), not complex
    int bytes_per = 16;
};

// Gather: out[e] = beta out[e] + alpha sum_t f_t(src_t[map]), map entries
// (index << 2) | (negate << 1) | conjugate, -1 for a zero term; term-major.
struct GatherOp {
    int out = -1;
    double alpha = 1.0, beta = 0.0;
    int T = 0;
    std::vector<int> src;      // distinct source arrays (at most 4)
    std::vector<int> slot;     // per term: its source among src
    int64_t n = 0;
    std::vector<int32_t> map;
};

// Bilinear gather: out[e] = c0 * f(src0[m0]) + sum_t sg_t A[p_t] B[q_t].
struct BilOp {
    int out = -1, src0 = -1, A = -1, B = -1;
    double c0 = 1.0;
    int np = 0;
    double sg[2] = {0, 0};
    int64_t n = 0;
    std::vector<int32_t> m0, p[2], q[2];
};

// One group of a grouped product: offsets in elements into the arrays.
struct PGroup {
    int m = 0, n = 0, k = 0;
    int A = -1, B = -1, C = -1;
    int64_t offA = 0, lda = 1, offB = 0, ldb = 1, offC = 0, ldc = 1;
    double alpha = 1.0, beta = 0.0;
};

// The K-split partial sums of one group: C = alpha sum_s W_s + beta C.
struct RedDesc {
    int C = -1;
    int64_t offC = 0, ldc = 1, offW = 0;
    int m = 0, n = 0, nsplit = 1;
    double alpha = 1.0, beta = 0.0;
};

struct ProdOp {
    int opa = 0, opb = 0, tag = 0;
    std::string name;
    std::vector<PGroup> g;          // the launch (split pieces write the work array)
    std::vector<RedDesc> red;       // K-split reductions (empty: none)
    int64_t ws_elems = 0;           // work array elements this product uses
    // kernel: 0 the 96 x 64 MFMA tile, 1 the 32 x 32 one, 2 the 128 x 32 one,
    // 3 the streaming kernel (op N N, n <= nj <= 16: one row per thread)
    int variant = 0, nj = 0;
    double flops = 0;               // real flops of the classical four-product form
};

// Element-wise operations on whole arrays (same layout).
enum { EW_COPY = 0, EW_CONJ = 1, EW_AXPBY = 2, EW_FILL = 3 };
struct EwOp {
    int kind = 0, x = -1, y = -1;
    double a = 1.0, b = 0.0;
};

struct DenomOp {
    int S1 = -1, S2 = -1;
    bool divide = true;       // divide by the denominators (DENOM)
    bool mask = true;         // zero the frozen-core entries (ZCORE)
};

enum { OK_PROD = 0, OK_GATHER = 1, OK_BIL = 2, OK_EW = 3, OK_DENOM = 4 };
struct Op {
    int kind = 0, idx = 0;
};

// Array ids of the program (constants, state, work).
enum ArrId {
    A_FOO, A_FVO, A_FVV, A_OOOO, A_VOOO, A_VVOO, A_VOVO, A_VOVV, A_VVVV,
    A_cE2V, A_cF1V, A_cRV, A_cRO, A_cQ, A_cM1, A_cMGO, A_cMH, A_cVOOOt, A_cVOVVt, A_cVOVOt,
    A_D1, A_D2, A_F1m, A_F2m,
    A_T, A_T1, A_T2, A_S, A_S1, A_S2, A_TAU,
    A_HO, A_HV, A_HOV, A_W, A_B2, A_GO, A_GV, A_AZ, A_X, A_E2, A_F1, A_RING, A_RING2, A_H,
    A_HAKCI, A_Z, A_YAZ, A_YPa, A_YPb, A_WC, A_Y1, A_WCt, A_ZX, A_Xt, A_ZX2, A_WS,
    A_COUNT
};

// Term sub-programs, in CCEQN's order (cceqns.F:293-569), and the others.
enum Prog {
    P_TAU, P_RING, P_X, P_HOINTM, P_HVINTM, P_HOVINTM, P_T1EQNS, P_GOINTM, P_GVINTM, P_AINTM,
    P_HINTM_HAKCI, P_HINTM, P_T2EQNS, P_BINTM, P_DENOM, P_ZCORE, P_TAUS, P_MP2, P_CONST,
    P_DIVIDE, P_COUNT
};

struct Plan {
    Ctx c;
    std::vector<Space> sp;
    std::vector<Arr> arr;
    std::vector<ProdOp> prods;
    std::vector<GatherOp> gathers;
    std::vector<BilOp> bils;
    std::vector<EwOp> ews;
    std::vector<DenomOp> denoms;
    std::vector<Op> prog[P_COUNT];
    bool any_frozen = false;
    int64_t ws_elems = 1;
    std::string err;
};

// ------------------------------------------------------------------ builders
namespace detail {

// f(lo, hi) over [0, n) in contiguous chunks on up to 16 host threads
template <class F>
void par_for(int64_t n, F f) {
    const int64_t hw = std::max<int64_t>(1, std::min<int64_t>(16, std::thread::hardware_concurrency()));
    const int64_t nt = std::min<int64_t>(hw, std::max<int64_t>(1, n / 4096));
    if (nt <= 1) {
        f((int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int64_t t = 0; t < nt; ++t)
        th.emplace_back([&, t] { f(n * t / nt, n * (t + 1) / nt); });
    for (auto& x : th) x.join();
}

enum SpId { SP_1, SP_O, SP_V, SP_OO, SP_VV, SP_VO, SP_VVO, SP_VOO2, SP_OOO, SP_VOV, SP_VOOO,
            SP_OVO, SP_OOF, SP_COUNT };

inline int64_t blk_index(const Plan& P, int a, int64_t r, int64_t s) {
    const Arr& A = P.arr[a];
    const Space& R = P.sp[A.R];
    const Space& S = P.sp[A.S];
    const int g = R.el[r];
    if (S.el[s] != g) return -1;
    return A.off[g] + R.pos[r] + R.cnt[g] * (int64_t)S.pos[s];
}

inline void add_arr(Plan& P, int id, const char* name, int R, int S, bool real = false) {
    Arr& A = P.arr[id];
    A.name = name;
    A.R = R;
    A.S = S;
    A.real = real;
    A.bytes_per = real ? 8 : 16;
    const int N = 2 * P.c.G.n;
    A.off.assign(N + 1, 0);
    for (int g = 0; g < N; ++g) A.off[g + 1] = A.off[g] + P.sp[R].cnt[g] * P.sp[S].cnt[g];
    A.size = A.off[N];
}

struct Term {
    int src = 0;       // term slot
    int64_t r = 0, s = 0;
    int sign = 0;      // 0: absent
    bool conj = false;
};

// Gather over the output's blocked elements; fn(r, s, terms) fills T terms
// with C1 row/column indices of the sources (in the sources' own spaces).
template <class F>
int gather(Plan& P, int out, int T, std::vector<int> srcs, double alpha, double beta, F fn) {
    GatherOp G;
    G.out = out;
    G.T = T;
    G.alpha = alpha;
    G.beta = beta;
    for (int t = 0; t < T; ++t) {
        int k = 0;
        while (k < (int)G.src.size() && G.src[k] != srcs[t]) ++k;
        if (k == (int)G.src.size()) G.src.push_back(srcs[t]);
        G.slot.push_back(k);
    }
    if (G.src.size() > 4) {
        P.err = "blocked plan: a gather with more than four source arrays";
        return -1;
    }
    const Arr& O = P.arr[out];
    G.n = O.size;
    G.map.assign((size_t)T * G.n, -1);
    const Space& R = P.sp[O.R];
    const Space& S = P.sp[O.S];
    for (int t = 0; t < T; ++t)
        if (P.arr[srcs[t]].size >= ((int64_t)1 << 29)) {
            P.err = "blocked plan: gather source beyond 2^29 elements";
            return -1;
        }
    const int N = 2 * P.c.G.n;
    for (int g = 0; g < N; ++g) {
        const int64_t nr = R.cnt[g], ns = S.cnt[g];
        par_for(nr * ns, [&](int64_t x0, int64_t x1) {
            std::vector<Term> tm(T);
            if (x0 >= x1) return;
            int64_t js = x0 / nr, ir = x0 - js * nr;
            for (int64_t x = x0; x < x1; ++x, ++ir) {
                if (ir == nr) {
                    ir = 0;
                    ++js;
                }
                const int64_t e = O.off[g] + x;
                for (int t = 0; t < T; ++t) tm[t] = Term{t, 0, 0, 0, false};
                fn((int64_t)R.list[g][ir], (int64_t)S.list[g][js], tm.data());
                for (int t = 0; t < T; ++t) {
                    if (tm[t].sign == 0) continue;
                    const int64_t b = blk_index(P, srcs[t], tm[t].r, tm[t].s);
                    if (b < 0) continue;
                    G.map[(size_t)t * G.n + e] =
                        (int32_t)((b << 2) | (tm[t].sign < 0 ? 2 : 0) | (tm[t].conj ? 1 : 0));
                }
            }
        });
    }
    P.gathers.push_back(std::move(G));
    return (int)P.gathers.size() - 1;
}

// op(X) of a stored block rows x cols: (m x k) for A, (k x n) for B
inline void opdims(int op, int64_t rows, int64_t cols, int64_t& d0, int64_t& d1) {
    if (op == 0 || op == 3) { d0 = rows; d1 = cols; }
    else { d0 = cols; d1 = rows; }
}

// Grouped product C = alpha op(A) op(B) + beta C over the blocks.  Views:
// the arrays' own, or aliases (space ids) for A / B / C given as (R, S).
struct View { int arr, R, S; };

inline int product(Plan& P, const char* name, int opa, int opb, View A, View B, View C,
                   double alpha, double beta, int tag = 0) {
    ProdOp op;
    op.name = name;
    op.opa = opa;
    op.opb = opb;
    op.tag = tag;
    const int N = 2 * P.c.G.n;
    auto off_of = [&](const View& V, int g) {
        // offset of block g of view V inside its array (the alias must have
        // the array's blocks in the same places: checked by the caller)
        const Space& R = P.sp[V.R];
        const Space& S = P.sp[V.S];
        (void)S;
        if (V.R == P.arr[V.arr].R && V.S == P.arr[V.arr].S) return P.arr[V.arr].off[g];
        int64_t off = 0;
        for (int h = 0; h < g; ++h) off += P.sp[V.R].cnt[h] * P.sp[V.S].cnt[h];
        (void)R;
        return off;
    };
    int64_t tiles = 0;
    struct Raw { int g; int64_t m, n, k; };
    std::vector<Raw> raw;
    for (int g = 0; g < N; ++g) {
        int64_t am, ak, bk, bn;
        opdims(opa, P.sp[A.R].cnt[g], P.sp[A.S].cnt[g], am, ak);
        opdims(opb, P.sp[B.R].cnt[g], P.sp[B.S].cnt[g], bk, bn);
        const int64_t m = P.sp[C.R].cnt[g], n = P.sp[C.S].cnt[g];
        if (am != m || bn != n || ak != bk) {
            P.err = std::string("blocked plan: product ") + name + " does not chain";
            return -1;
        }
        if (m == 0 || n == 0) continue;
        // a block with k = 0 stays a group: the kernels then write beta C
        if (ak == 0 && beta == 1.0) continue;
        raw.push_back({g, m, n, ak});
        tiles += ((m + 95) / 96) * ((n + 63) / 64);
        op.flops += 8.0 * (double)m * (double)n * (double)ak;
    }
    // the kernel: streaming for N N products with an occupied-spinor width
    // (n <= 16) and a large A; the 32 x 32 tile for small blocks; otherwise the
    // 96 x 64 or the 128 x 32 tile, whichever computes less padding
    int64_t mx = 0, nx = 0, rowblocks = 0;
    double useful = 0, pad96 = 0, pad128 = 0;
    for (const Raw& r : raw) {
        mx = std::max(mx, r.m);
        nx = std::max(nx, r.n);
        useful += (double)r.m * r.n * r.k;
        pad96 += (double)((r.m + 95) / 96 * 96) * ((r.n + 63) / 64 * 64) * r.k;
        pad128 += (double)((r.m + 127) / 128 * 128) * ((r.n + 31) / 32 * 32) * r.k;
        rowblocks += (r.m + 255) / 256;
    }
    // (the streaming kernel has no K split: it needs rows for every CU)
    if (opa == 0 && opb == 0 && nx <= 16 && rowblocks >= 256) {
        op.variant = 3;
        op.nj = nx <= 4 ? 4 : nx <= 8 ? 8 : 16;
    } else if (mx <= 48 || tiles < 64) {
        op.variant = 1;
    } else if (nx <= 32 || (useful > 0 && useful / pad128 > useful / pad96 + 0.05)) {
        op.variant = 2;
    } else {
        op.variant = 0;
    }
    // K-split: each group's k cut into chunks, a split count s shared out in
    // proportion to the groups' k, chosen to minimise the makespan bound
    // max(longest workgroup, total work / workgroup slots), with a fixed cost
    // per workgroup (in k units) for its prologue and epilogue
    const int64_t slots = op.variant == 1 ? 2048 : 512, ovh = 64;
    int64_t kmax = 1;
    for (const Raw& r : raw) kmax = std::max(kmax, r.k);
    auto tiles_of = [&](const Raw& r) {
        return op.variant == 1 ? ((r.m + 31) / 32) * ((r.n + 31) / 32)
               : op.variant == 2 ? ((r.m + 127) / 128) * ((r.n + 31) / 32)
                                 : ((r.m + 95) / 96) * ((r.n + 63) / 64);
    };
    auto splits_for = [&](int sc, std::vector<int64_t>& chunk) {
        chunk.assign(raw.size(), 0);
        double work = 0.0;
        int64_t longest = 0;
        for (size_t i = 0; i < raw.size(); ++i) {
            const Raw& r = raw[i];
            int64_t si = std::max<int64_t>(1, (sc * r.k + kmax / 2) / kmax);
            si = std::min<int64_t>(si, std::max<int64_t>(1, r.k / 256));
            int64_t c = (r.k + si - 1) / si;
            if (si > 1) c = (c + 15) / 16 * 16;
            chunk[i] = std::max<int64_t>(c, 1);
            si = (r.k + chunk[i] - 1) / chunk[i];
            work += (double)tiles_of(r) * si * (chunk[i] + ovh);
            longest = std::max(longest, chunk[i] + ovh);
        }
        return std::max((double)longest, work / slots);
    };
    std::vector<int64_t> chunks, trial;
    double best = splits_for(1, chunks);
    // (up to 1024 pieces: products with a tiny output and a very long k, such
    // as HOINTM's o_g x o_g from k = |(VV, o)_g|, are latency-bound per piece)
    for (int sc = 2; sc <= 1024 && op.variant != 3; sc += sc < 64 ? 1 : 16) {
        const double c = splits_for(sc, trial);
        if (c < 0.97 * best) {
            best = c;
            chunks = trial;
        }
    }
    int64_t wsoff = 0;
    for (size_t ri = 0; ri < raw.size(); ++ri) {
        const Raw& r = raw[ri];
        int nsplit = r.k > 0 ? (int)((r.k + chunks[ri] - 1) / chunks[ri]) : 1;
        const bool AR = opa == 0 || opa == 3;     // A stored m-fastest
        const bool BR = !(opb == 0 || opb == 3);  // B stored n-fastest
        const int64_t lda = std::max<int64_t>(1, P.sp[A.R].cnt[r.g]);
        const int64_t ldb = std::max<int64_t>(1, P.sp[B.R].cnt[r.g]);
        const int64_t ldc = std::max<int64_t>(1, r.m);
        const int64_t oA = off_of(A, r.g), oB = off_of(B, r.g), oC = off_of(C, r.g);
        if (nsplit == 1) {
            PGroup pg;
            pg.m = (int)r.m; pg.n = (int)r.n; pg.k = (int)r.k;
            pg.A = A.arr; pg.B = B.arr; pg.C = C.arr;
            pg.offA = oA; pg.lda = lda; pg.offB = oB; pg.ldb = ldb; pg.offC = oC; pg.ldc = ldc;
            pg.alpha = alpha;
            pg.beta = beta;
            op.g.push_back(pg);
            continue;
        }
        const int64_t chunk = chunks[ri];
        RedDesc rd;
        rd.C = C.arr; rd.offC = oC; rd.ldc = ldc; rd.offW = wsoff;
        rd.m = (int)r.m; rd.n = (int)r.n; rd.nsplit = nsplit;
        rd.alpha = alpha; rd.beta = beta;
        for (int s = 0; s < nsplit; ++s) {
            const int64_t kb = s * chunk, kl = std::min(chunk, r.k - kb);
            PGroup pg;
            pg.m = (int)r.m; pg.n = (int)r.n; pg.k = (int)kl;
            pg.A = A.arr; pg.B = B.arr; pg.C = A_WS;
            pg.offA = oA + (AR ? kb * lda : kb);
            pg.lda = lda;
            pg.offB = oB + (BR ? kb * ldb : kb);
            pg.ldb = ldb;
            pg.offC = wsoff + s * r.m * r.n;
            pg.ldc = r.m;
            pg.alpha = 1.0;
            pg.beta = 0.0;
            op.g.push_back(pg);
        }
        wsoff += (int64_t)nsplit * r.m * r.n;
        op.red.push_back(rd);
    }
    op.ws_elems = wsoff;
    P.ws_elems = std::max(P.ws_elems, wsoff);
    P.prods.push_back(std::move(op));
    return (int)P.prods.size() - 1;
}

inline int ew(Plan& P, int kind, int x, int y, double a = 1.0, double b = 0.0) {
    EwOp e;
    e.kind = kind;
    e.x = x;
    e.y = y;
    e.a = a;
    e.b = b;
    P.ews.push_back(e);
    return (int)P.ews.size() - 1;
}

// an alias view (R, S) of array a must put every element where the array's
// own view does
inline bool alias_ok(const Plan& P, int a, int R, int S) {
    const Arr& A = P.arr[a];
    const Space& r0 = P.sp[A.R];
    const Space& s0 = P.sp[A.S];
    const Space& r1 = P.sp[R];
    const Space& s1 = P.sp[S];
    if (r0.n1 * s0.n1 != r1.n1 * s1.n1) return false;
    int64_t off1 = 0;
    const int N = 2 * P.c.G.n;
    std::vector<int64_t> o1(N + 1, 0);
    for (int g = 0; g < N; ++g) o1[g + 1] = o1[g] + r1.cnt[g] * s1.cnt[g];
    if (o1[N] != A.size) return false;
    (void)off1;
    for (int64_t x = 0; x < r0.n1 * s0.n1; ++x) {
        const int64_t ra = x % r0.n1, sa = x / r0.n1;
        const int64_t rb = x % r1.n1, sb = x / r1.n1;
        const int g0 = r0.el[ra], g1 = r1.el[rb];
        const bool nz0 = s0.el[sa] == g0, nz1 = s1.el[sb] == g1;
        if (nz0 != nz1) return false;
        if (!nz0) continue;
        const int64_t b0 = A.off[g0] + r0.pos[ra] + r0.cnt[g0] * (int64_t)s0.pos[sa];
        const int64_t b1 = o1[g1] + r1.pos[rb] + r1.cnt[g1] * (int64_t)s1.pos[sb];
        if (b0 != b1) return false;
    }
    return true;
}

}  // namespace detail

// Builds the whole plan: spaces, arrays, the iteration's operations, the
// constant builders, MP2 and the denominators.
inline int build_plan(const SymIn& s, Plan& P) {
    using namespace detail;
    if ((int)s.ptab.size() != s.nrep * s.nrep) {
        P.err = "blocked plan: bad pair table";
        return -1;
    }
    if (!make_group(s, P.c.G)) {
        P.err = "blocked plan: irreps are not self-inverse";
        return -1;
    }
    make_ctx(s, P.c.G, P.c);
    Ctx& c = P.c;
    const int o = c.o, v = c.v;
    for (int f : s.frozen) P.any_frozen = P.any_frozen || f != 0;
    P.sp.assign(SP_COUNT, Space());
    make_space(c, {}, P.sp[SP_1]);
    make_space(c, {AT_O}, P.sp[SP_O]);
    make_space(c, {AT_V}, P.sp[SP_V]);
    make_space(c, {AT_OO}, P.sp[SP_OO]);
    make_space(c, {AT_VV}, P.sp[SP_VV]);
    make_space(c, {AT_V, AT_O}, P.sp[SP_VO]);
    make_space(c, {AT_VV, AT_O}, P.sp[SP_VVO]);
    make_space(c, {AT_V, AT_OO}, P.sp[SP_VOO2]);
    make_space(c, {AT_OO, AT_O}, P.sp[SP_OOO]);
    make_space(c, {AT_V, AT_O, AT_V}, P.sp[SP_VOV]);
    make_space(c, {AT_V, AT_O, AT_O}, P.sp[SP_VOOO]);
    make_space(c, {AT_O, AT_V, AT_O}, P.sp[SP_OVO]);
    make_space(c, {AT_O, AT_O}, P.sp[SP_OOF]);
    P.arr.assign(A_COUNT, Arr());
    add_arr(P, A_FOO, "FOO", SP_O, SP_O);
    add_arr(P, A_FVO, "FVO", SP_V, SP_O);
    add_arr(P, A_FVV, "FVV", SP_V, SP_V);
    add_arr(P, A_OOOO, "OOOO", SP_OO, SP_OO);
    add_arr(P, A_VOOO, "VOOO", SP_VO, SP_OO);
    add_arr(P, A_VVOO, "VVOO", SP_VV, SP_OO);
    add_arr(P, A_VOVO, "VOVO", SP_VO, SP_VO);
    add_arr(P, A_VOVV, "VOVV", SP_VO, SP_VV);
    add_arr(P, A_VVVV, "VVVV", SP_VV, SP_VV);
    add_arr(P, A_cE2V, "VVOO_E2", SP_VVO, SP_O);
    add_arr(P, A_cF1V, "VVOO_F1", SP_V, SP_VOO2);
    add_arr(P, A_cRV, "VVOO_R", SP_VO, SP_VO);
    add_arr(P, A_cRO, "VOVO_R", SP_VO, SP_VO);
    add_arr(P, A_cQ, "VOVV_Q", SP_VOV, SP_V);
    add_arr(P, A_cM1, "M1", SP_VOO2, SP_O);
    add_arr(P, A_cMGO, "MGO", SP_VO, SP_OOF);
    add_arr(P, A_cMH, "MH", SP_O, SP_OVO);
    add_arr(P, A_cVOOOt, "VOOOt", SP_OOO, SP_V);
    add_arr(P, A_cVOVVt, "VOVVt", SP_VVO, SP_V);
    add_arr(P, A_cVOVOt, "VOVOt", SP_VOOO, SP_V);
    add_arr(P, A_D1, "D1", SP_V, SP_O, true);
    add_arr(P, A_D2, "D2", SP_VV, SP_OO, true);
    add_arr(P, A_F1m, "F1m", SP_V, SP_O, true);
    add_arr(P, A_F2m, "F2m", SP_VV, SP_OO, true);
    add_arr(P, A_T1, "T1", SP_V, SP_O);
    add_arr(P, A_T2, "T2", SP_VV, SP_OO);
    add_arr(P, A_S1, "S1", SP_V, SP_O);
    add_arr(P, A_S2, "S2", SP_VV, SP_OO);
    // T = [T1 | T2] and S = [S1 | S2] are the amplitude vectors (views over
    // the same memory: T1 / S1 at offset 0, T2 / S2 at offset |T1|)
    P.arr[A_T] = P.arr[A_T1];
    P.arr[A_T].name = "T";
    P.arr[A_T].size = P.arr[A_T1].size + P.arr[A_T2].size;
    P.arr[A_S] = P.arr[A_T];
    P.arr[A_S].name = "S";
    add_arr(P, A_TAU, "TAU", SP_VV, SP_OO);
    add_arr(P, A_HO, "HO", SP_O, SP_O);
    add_arr(P, A_HV, "HV", SP_V, SP_V);
    add_arr(P, A_HOV, "HOV", SP_V, SP_O);
    add_arr(P, A_W, "W", SP_V, SP_O);
    add_arr(P, A_B2, "B2", SP_O, SP_O);
    add_arr(P, A_GO, "GO", SP_O, SP_O);
    add_arr(P, A_GV, "GV", SP_V, SP_V);
    add_arr(P, A_AZ, "AZ", SP_OO, SP_OO);
    add_arr(P, A_X, "X", SP_VO, SP_OO);
    add_arr(P, A_E2, "E2", SP_VVO, SP_O);
    add_arr(P, A_F1, "F1", SP_V, SP_VOO2);
    add_arr(P, A_RING, "RING", SP_VO, SP_VO);
    add_arr(P, A_RING2, "RING2", SP_VO, SP_VO);
    add_arr(P, A_H, "H", SP_VO, SP_VO);
    add_arr(P, A_HAKCI, "HAKCI", SP_VOV, SP_O);
    add_arr(P, A_Z, "Z", SP_V, SP_OVO);
    add_arr(P, A_YAZ, "YAZ", SP_OOO, SP_O);
    add_arr(P, A_YPa, "YPa", SP_VVO, SP_O);
    add_arr(P, A_YPb, "YPb", SP_VVO, SP_O);
    add_arr(P, A_WC, "WC", SP_VO, SP_OO);
    add_arr(P, A_Y1, "Y1", SP_VOOO, SP_O);
    add_arr(P, A_WCt, "WCt", SP_O, SP_VOO2);
    add_arr(P, A_ZX, "ZX", SP_V, SP_VOO2);
    add_arr(P, A_Xt, "Xt", SP_VOO2, SP_O);
    add_arr(P, A_ZX2, "ZX2", SP_VOO2, SP_V);
    P.arr[A_WS].name = "WS";
    // the [VO ; 1] views of the VO vectors and the [oo ; 1] view of GO are
    // their own blocked layouts
    if (!alias_ok(P, A_T1, SP_VO, SP_1) || !alias_ok(P, A_GO, SP_OOF, SP_1)) {
        P.err = "blocked plan: vector views do not coincide";
        return -1;
    }
    const int64_t nvo = (int64_t)o * v;
    (void)nvo;
    auto OO = [&](int x, int y) { return pair_ix(x, y, o); };
    auto VV = [&](int x, int y) { return pair_ix(x, y, v); };
    auto& pg = P.prog;
    auto addp = [&](int prog, int kind, int idx) -> int {
        if (idx < 0) return -1;
        pg[prog].push_back(Op{kind, idx});
        return 0;
    };
#define RCC_ADD(prog, kind, call)            \
    do {                                     \
        if (addp(prog, kind, (call)) != 0) { \
            if (P.err.empty()) P.err = "blocked plan: build failed"; \
            return -1;                       \
        }                                    \
    } while (0)

    // ---------------------------------------------------- constants
    // VVOO_E2[(AB, y), x] = +-VVOO[AB, {x, y}]  (expand_e2: last pair expanded)
    RCC_ADD(P_CONST, OK_GATHER, gather(P, A_cE2V, 1, {A_VVOO}, 1.0, 0.0,
        [&](int64_t r, int64_t x, Term* t) {
            const int64_t ab = r % c.nvvt;
            const int y = (int)(r / c.nvvt);
            if (x == y) return;
            t[0].r = ab;
            t[0].s = x > y ? OO((int)x, y) : OO(y, (int)x);
            t[0].sign = x > y ? 1 : -1;
        }));
    // VVOO_F1[a, (d, col)] = +-VVOO[{a, d}, col]  (expand_f1: first pair expanded)
    RCC_ADD(P_CONST, OK_GATHER, gather(P, A_cF1V, 1, {A_VVOO}, 1.0, 0.0,
        [&](int64_t a, int64_t sc, Term* t) {
            const int d = (int)(sc % v);
            const int64_t col = sc / v;
            if (a == d) return;
            t[0].r = a > d ? VV((int)a, d) : VV(d, (int)a);
            t[0].s = col;
            t[0].sign = a > d ? 1 : -1;
        }));
    // ring form of a [VV ; OO] array: out[(a,i),(c,k)] = +-P[{a,c},{i,k}]
    auto ring_of = [&](int64_t r, int64_t sc, Term* t) {
        const int a = (int)(r % v), i = (int)(r / v);
        const int cc = (int)(sc % v), k = (int)(sc / v);
        if (a == cc || i == k) return;
        t[0].r = a > cc ? VV(a, cc) : VV(cc, a);
        t[0].s = i > k ? OO(i, k) : OO(k, i);
        t[0].sign = ((a > cc) == (i > k)) ? 1 : -1;
    };
    RCC_ADD(P_CONST, OK_GATHER, gather(P, A_cRV, 1, {A_VVOO}, 1.0, 0.0, ring_of));
    // VOVO_R[(a,i),(c,k)] = VOVO[(a,k),(c,i)]  (perm_akci, SRT16)
    auto akci = [&](int64_t r, int64_t sc, Term* t) {
        const int a = (int)(r % v), i = (int)(r / v);
        const int cc = (int)(sc % v), k = (int)(sc / v);
        t[0].r = a + (int64_t)v * k;
        t[0].s = cc + (int64_t)v * i;
        t[0].sign = 1;
    };
    RCC_ADD(P_CONST, OK_GATHER, gather(P, A_cRO, 1, {A_VOVO}, 1.0, 0.0, akci));
    // VOVV_Q[((a,k), c), d] = +-VOVV[(a,k), {c, d}]
    RCC_ADD(P_CONST, OK_GATHER, gather(P, A_cQ, 1, {A_VOVV}, 1.0, 0.0,
        [&](int64_t r, int64_t d, Term* t) {
            const int64_t ak = r % ((int64_t)v * o);
            const int cc = (int)(r / ((int64_t)v * o));
            if (cc == d) return;
            t[0].r = ak;
            t[0].s = cc > d ? VV(cc, (int)d) : VV((int)d, cc);
            t[0].sign = cc > d ? 1 : -1;
        }));
    // M1[(c, kl), i] = conj VOOO[(c, i), kl]
    RCC_ADD(P_CONST, OK_GATHER, gather(P, A_cM1, 1, {A_VOOO}, 1.0, 0.0,
        [&](int64_t r, int64_t i, Term* t) {
            const int cc = (int)(r % v);
            const int64_t kl = r / v;
            t[0].r = cc + (int64_t)v * i;
            t[0].s = kl;
            t[0].sign = 1;
            t[0].conj = true;
        }));
    // MGO[(c, l), (k, i)] = +-conj VOOO[(c, i), {k, l}]
    RCC_ADD(P_CONST, OK_GATHER, gather(P, A_cMGO, 1, {A_VOOO}, 1.0, 0.0,
        [&](int64_t r, int64_t sc, Term* t) {
            const int cc = (int)(r % v), l = (int)(r / v);
            const int k = (int)(sc % o), i = (int)(sc / o);
            if (k == l) return;
            t[0].r = cc + (int64_t)v * i;
            t[0].s = k > l ? OO(k, l) : OO(l, k);
            t[0].sign = k > l ? 1 : -1;
            t[0].conj = true;
        }));
    // MH[l, (k, c, i)] = +-conj VOOO[(c, i), {l, k}]
    RCC_ADD(P_CONST, OK_GATHER, gather(P, A_cMH, 1, {A_VOOO}, 1.0, 0.0,
        [&](int64_t l, int64_t sc, Term* t) {
            const int k = (int)(sc % o);
            const int64_t ci = sc / o;
            if (k == l) return;
            t[0].r = ci;
            t[0].s = l > k ? OO((int)l, k) : OO(k, (int)l);
            t[0].sign = l > k ? 1 : -1;
            t[0].conj = true;
        }));
    // VOOOt[(kl, i), c] = conj VOOO[(c, i), kl]
    RCC_ADD(P_CONST, OK_GATHER, gather(P, A_cVOOOt, 1, {A_VOOO}, 1.0, 0.0,
        [&](int64_t r, int64_t cc, Term* t) {
            const int64_t kl = r % c.noot;
            const int i = (int)(r / c.noot);
            t[0].r = cc + (int64_t)v * i;
            t[0].s = kl;
            t[0].sign = 1;
            t[0].conj = true;
        }));
    // VOVVt[(AB, j), c] = conj VOVV[(c, j), AB]
    RCC_ADD(P_CONST, OK_GATHER, gather(P, A_cVOVVt, 1, {A_VOVV}, 1.0, 0.0,
        [&](int64_t r, int64_t cc, Term* t) {
            const int64_t ab = r % c.nvvt;
            const int j = (int)(r / c.nvvt);
            t[0].r = cc + (int64_t)v * j;
            t[0].s = ab;
            t[0].sign = 1;
            t[0].conj = true;
        }));
    // VOVOt[((b,k), i), c] = VOVO[(b,k), (c,i)]
    RCC_ADD(P_CONST, OK_GATHER, gather(P, A_cVOVOt, 1, {A_VOVO}, 1.0, 0.0,
        [&](int64_t r, int64_t cc, Term* t) {
            const int64_t bk = r % ((int64_t)v * o);
            const int i = (int)(r / ((int64_t)v * o));
            t[0].r = bk;
            t[0].s = cc + (int64_t)v * i;
            t[0].sign = 1;
        }));

    // ---------------------------------------------------- amplitude-derived
    // tau = T2 + t(a,i) t(b,j) - t(a,j) t(b,i)  (GETTAU ccintm.F:833)
    auto tau_bil = [&](int out, int t1src, int t2src) -> int {
        BilOp B;
        B.out = out;
        B.src0 = t2src;
        B.A = t1src;
        B.B = t1src;
        B.c0 = 1.0;
        B.np = 2;
        B.sg[0] = 1.0;
        B.sg[1] = -1.0;
        const Arr& O = P.arr[out];
        B.n = O.size;
        B.m0.assign(B.n, -1);
        for (int q = 0; q < 2; ++q) {
            B.p[q].assign(B.n, -1);
            B.q[q].assign(B.n, -1);
        }
        const Space& R = P.sp[O.R];
        const Space& S = P.sp[O.S];
        const int N = 2 * c.G.n;
        const int t1arr = t1src == A_S ? A_S1 : A_T1;
        const int t2arr = t2src == A_S ? A_S2 : A_T2;
        const int64_t t2base = t2src == A_S || t2src == A_T ? P.arr[A_T1].size : 0;
        if (P.arr[A_T].size >= ((int64_t)1 << 29)) {
            P.err = "blocked plan: amplitudes beyond 2^29 elements";
            return -1;
        }
        for (int g = 0; g < N; ++g)
            par_for(S.cnt[g], [&](int64_t j0, int64_t j1) {
            for (int64_t js = j0; js < j1; ++js)
                for (int64_t ir = 0; ir < R.cnt[g]; ++ir) {
                    const int64_t e = O.off[g] + ir + R.cnt[g] * js;
                    const int64_t ab = R.list[g][ir], ij = S.list[g][js];
                    const int a = c.vvx[ab], b = c.vvy[ab], i = c.oox[ij], j = c.ooy[ij];
                    B.m0[e] = (int32_t)((t2base + blk_index(P, t2arr, ab, ij)) << 2);
                    const int64_t ai = blk_index(P, t1arr, a, i), bj = blk_index(P, t1arr, b, j);
                    const int64_t aj = blk_index(P, t1arr, a, j), bi = blk_index(P, t1arr, b, i);
                    if (ai >= 0 && bj >= 0) {
                        B.p[0][e] = (int32_t)ai;
                        B.q[0][e] = (int32_t)bj;
                    }
                    if (aj >= 0 && bi >= 0) {
                        B.p[1][e] = (int32_t)aj;
                        B.q[1][e] = (int32_t)bi;
                    }
                }
            });
        P.bils.push_back(std::move(B));
        return (int)P.bils.size() - 1;
    };
    // the T1 / T2 sources of tau index into T (or S): T2 at offset |T1|
    RCC_ADD(P_TAU, OK_BIL, tau_bil(A_TAU, A_T, A_T));
    RCC_ADD(P_TAUS, OK_BIL, tau_bil(A_TAU, A_S, A_S));
    // RING[(a,i),(c,k)] = +-T2[{a,c},{i,k}]  (expand_ring, s = 1, no T1)
    RCC_ADD(P_RING, OK_GATHER, gather(P, A_RING, 1, {A_T2}, 1.0, 0.0, ring_of));
    // X = VOVV . TAU  (ccintm.F:772-788)
    RCC_ADD(P_X, OK_PROD, product(P, "X", 0, 0, {A_VOVV, SP_VO, SP_VV}, {A_TAU, SP_VV, SP_OO},
                                  {A_X, SP_VO, SP_OO}, 1.0, 0.0));

    // ---------------------------------------------------- HOINTM
    auto offdiag = [&](int out, int src) {
        return gather(P, out, 1, {src}, 1.0, 0.0, [&](int64_t r, int64_t sc, Term* t) {
            if (r == sc) return;
            t[0].r = r;
            t[0].s = sc;
            t[0].sign = 1;
        });
    };
    // E2[(AB, y), x] = +-P[AB, {x, y}]
    auto expand_e2 = [&](int src) {
        return gather(P, A_E2, 1, {src}, 1.0, 0.0, [&](int64_t r, int64_t x, Term* t) {
            const int64_t ab = r % c.nvvt;
            const int y = (int)(r / c.nvvt);
            if (x == y) return;
            t[0].r = ab;
            t[0].s = x > y ? OO((int)x, y) : OO(y, (int)x);
            t[0].sign = x > y ? 1 : -1;
        });
    };
    // F1[a, (d, col)] = +-P[{a, d}, col]
    auto expand_f1 = [&](int src) {
        return gather(P, A_F1, 1, {src}, 1.0, 0.0, [&](int64_t a, int64_t sc, Term* t) {
            const int d = (int)(sc % v);
            const int64_t col = sc / v;
            if (a == d) return;
            t[0].r = a > d ? VV((int)a, d) : VV(d, (int)a);
            t[0].s = col;
            t[0].sign = a > d ? 1 : -1;
        });
    };
    RCC_ADD(P_HOINTM, OK_GATHER, offdiag(A_HO, A_FOO));
    RCC_ADD(P_HOINTM, OK_GATHER, expand_e2(A_TAU));
    RCC_ADD(P_HOINTM, OK_PROD, product(P, "HOINTM", 2, 0, {A_cE2V, SP_VVO, SP_O},
                                       {A_E2, SP_VVO, SP_O}, {A_HO, SP_O, SP_O}, 1.0, 1.0));
    // ---------------------------------------------------- HVINTM
    RCC_ADD(P_HVINTM, OK_GATHER, offdiag(A_HV, A_FVV));
    RCC_ADD(P_HVINTM, OK_GATHER, expand_f1(A_TAU));
    RCC_ADD(P_HVINTM, OK_PROD, product(P, "HVINTM", 0, 2, {A_F1, SP_V, SP_VOO2},
                                       {A_cF1V, SP_V, SP_VOO2}, {A_HV, SP_V, SP_V}, -1.0, 1.0));
    // ---------------------------------------------------- HOVINTM
    RCC_ADD(P_HOVINTM, OK_EW, ew(P, EW_CONJ, A_FVO, A_HOV));
    RCC_ADD(P_HOVINTM, OK_PROD, product(P, "HOVINTM", 2, 0, {A_cRV, SP_VO, SP_VO},
                                        {A_T1, SP_VO, SP_1}, {A_HOV, SP_VO, SP_1}, 1.0, 1.0));
    // ---------------------------------------------------- T1EQNS
    RCC_ADD(P_T1EQNS, OK_EW, ew(P, EW_COPY, A_FVO, A_S1));
    RCC_ADD(P_T1EQNS, OK_EW, ew(P, EW_CONJ, A_FVO, A_W));
    RCC_ADD(P_T1EQNS, OK_EW, ew(P, EW_AXPBY, A_HOV, A_W, 1.0, -2.0));
    RCC_ADD(P_T1EQNS, OK_PROD, product(P, "T1EQNS_B2", 1, 0, {A_W, SP_V, SP_O},
                                       {A_T1, SP_V, SP_O}, {A_B2, SP_O, SP_O}, 1.0, 0.0));
    RCC_ADD(P_T1EQNS, OK_PROD, product(P, "T1EQNS_TB2", 0, 0, {A_T1, SP_V, SP_O},
                                       {A_B2, SP_O, SP_O}, {A_S1, SP_V, SP_O}, 1.0, 1.0));
    RCC_ADD(P_T1EQNS, OK_PROD, product(P, "T1EQNS_RING", 0, 0, {A_RING, SP_VO, SP_VO},
                                       {A_HOV, SP_VO, SP_1}, {A_S1, SP_VO, SP_1}, 1.0, 1.0));
    RCC_ADD(P_T1EQNS, OK_PROD, product(P, "T1EQNS_HV", 0, 0, {A_HV, SP_V, SP_V},
                                       {A_T1, SP_V, SP_O}, {A_S1, SP_V, SP_O}, 1.0, 1.0));
    RCC_ADD(P_T1EQNS, OK_PROD, product(P, "T1EQNS_HO", 0, 0, {A_T1, SP_V, SP_O},
                                       {A_HO, SP_O, SP_O}, {A_S1, SP_V, SP_O}, -1.0, 1.0));
    RCC_ADD(P_T1EQNS, OK_GATHER, expand_f1(A_TAU));
    RCC_ADD(P_T1EQNS, OK_PROD, product(P, "T1EQN1", 0, 0, {A_F1, SP_V, SP_VOO2},
                                       {A_cM1, SP_VOO2, SP_O}, {A_S1, SP_V, SP_O}, 1.0, 1.0));
    // T1EQN2: S1[a, i] += sum_k +-X[(a, k), {i, k}]
    RCC_ADD(P_T1EQNS, OK_GATHER, gather(P, A_S1, o, std::vector<int>(o, A_X), 1.0, 1.0,
        [&](int64_t a, int64_t i, Term* t) {
            for (int k = 0; k < o; ++k) {
                if (k == i) continue;
                t[k].r = a + (int64_t)v * k;
                t[k].s = i > k ? OO((int)i, k) : OO(k, (int)i);
                t[k].sign = i > k ? 1 : -1;
            }
        }));
    RCC_ADD(P_T1EQNS, OK_PROD, product(P, "T1EQNS_W", 0, 0, {A_cRO, SP_VO, SP_VO},
                                       {A_T1, SP_VO, SP_1}, {A_S1, SP_VO, SP_1}, -1.0, 1.0));
    // ---------------------------------------------------- GOINTM
    RCC_ADD(P_GOINTM, OK_EW, ew(P, EW_COPY, A_HO, A_GO));
    RCC_ADD(P_GOINTM, OK_PROD, product(P, "GOINTM_F", 2, 0, {A_FVO, SP_V, SP_O},
                                       {A_T1, SP_V, SP_O}, {A_GO, SP_O, SP_O}, 1.0, 1.0));
    RCC_ADD(P_GOINTM, OK_PROD, product(P, "GOINTM_V", 1, 0, {A_cMGO, SP_VO, SP_OOF},
                                       {A_T1, SP_VO, SP_1}, {A_GO, SP_OOF, SP_1}, -1.0, 1.0));
    // ---------------------------------------------------- GVINTM
    RCC_ADD(P_GVINTM, OK_EW, ew(P, EW_COPY, A_HV, A_GV));
    RCC_ADD(P_GVINTM, OK_PROD, product(P, "GVINTM_F", 0, 2, {A_T1, SP_V, SP_O},
                                       {A_FVO, SP_V, SP_O}, {A_GV, SP_V, SP_V}, -1.0, 1.0));
    RCC_ADD(P_GVINTM, OK_PROD, product(P, "HAKCI", 0, 0, {A_cQ, SP_VOV, SP_V},
                                       {A_T1, SP_V, SP_O}, {A_HAKCI, SP_VOV, SP_O}, -1.0, 0.0));
    // GV[a, c] += -sum_k HAKCI[(a, k, c), k]
    RCC_ADD(P_GVINTM, OK_GATHER, gather(P, A_GV, o, std::vector<int>(o, A_HAKCI), -1.0, 1.0,
        [&](int64_t a, int64_t cc, Term* t) {
            for (int k = 0; k < o; ++k) {
                t[k].r = a + (int64_t)v * k + (int64_t)v * o * cc;
                t[k].s = k;
                t[k].sign = 1;
            }
        }));
    // ---------------------------------------------------- AINTM
    RCC_ADD(P_AINTM, OK_EW, ew(P, EW_COPY, A_OOOO, A_AZ));
    if (o > 0 && c.noot > 0) {
        RCC_ADD(P_AINTM, OK_PROD, product(P, "AINTM_Y", 0, 0, {A_cVOOOt, SP_OOO, SP_V},
                                          {A_T1, SP_V, SP_O}, {A_YAZ, SP_OOO, SP_O}, 1.0, 0.0));
        // AZ[kl, {x > y}] += YAZ[(kl, y), x] - YAZ[(kl, x), y]
        RCC_ADD(P_AINTM, OK_GATHER, gather(P, A_AZ, 2, {A_YAZ, A_YAZ}, 1.0, 1.0,
            [&](int64_t kl, int64_t xy, Term* t) {
                const int x = c.oox[xy], y = c.ooy[xy];
                t[0].r = kl + c.noot * y; t[0].s = x; t[0].sign = 1;
                t[1].r = kl + c.noot * x; t[1].s = y; t[1].sign = -1;
            }));
    }
    RCC_ADD(P_AINTM, OK_PROD, product(P, "AINTM_TAU", 2, 0, {A_VVOO, SP_VV, SP_OO},
                                      {A_TAU, SP_VV, SP_OO}, {A_AZ, SP_OO, SP_OO}, 1.0, 1.0));
    // ---------------------------------------------------- HINTM
    RCC_ADD(P_HINTM_HAKCI, OK_PROD, product(P, "HAKCI", 0, 0, {A_cQ, SP_VOV, SP_V},
                                            {A_T1, SP_V, SP_O}, {A_HAKCI, SP_VOV, SP_O}, -1.0,
                                            0.0));
    RCC_ADD(P_HINTM, OK_PROD, product(P, "HINTM_MH", 0, 0, {A_T1, SP_V, SP_O},
                                      {A_cMH, SP_O, SP_OVO}, {A_Z, SP_V, SP_OVO}, 1.0, 0.0));
    // H[(a,i),(c,k)] = HAKCI[((a,k),c), i] - VOVO[(a,k),(c,i)] + Z[a, (k,c,i)]
    RCC_ADD(P_HINTM, OK_GATHER, gather(P, A_H, 3, {A_HAKCI, A_VOVO, A_Z}, 1.0, 0.0,
        [&](int64_t r, int64_t sc, Term* t) {
            const int a = (int)(r % v), i = (int)(r / v);
            const int cc = (int)(sc % v), k = (int)(sc / v);
            t[0].r = a + (int64_t)v * k + (int64_t)v * o * cc; t[0].s = i; t[0].sign = 1;
            t[1].r = a + (int64_t)v * k; t[1].s = cc + (int64_t)v * i; t[1].sign = -1;
            t[2].r = a; t[2].s = k + (int64_t)o * (cc + (int64_t)v * i); t[2].sign = 1;
        }));
    // RING2[(a,i),(c,k)] = 1/2 +-T2[{a,c},{i,k}] - t(a,k) t(c,i)
    {
        BilOp B;
        B.out = A_RING2;
        B.src0 = A_T;
        B.A = A_T;
        B.B = A_T;
        B.c0 = 0.5;
        B.np = 1;
        B.sg[0] = -1.0;
        const Arr& O = P.arr[A_RING2];
        B.n = O.size;
        B.m0.assign(B.n, -1);
        B.p[0].assign(B.n, -1);
        B.q[0].assign(B.n, -1);
        const Space& R = P.sp[O.R];
        const Space& S = P.sp[O.S];
        const int64_t t2base = P.arr[A_T1].size;
        const int N = 2 * c.G.n;
        for (int g = 0; g < N; ++g)
            par_for(S.cnt[g], [&](int64_t j0, int64_t j1) {
            for (int64_t js = j0; js < j1; ++js)
                for (int64_t ir = 0; ir < R.cnt[g]; ++ir) {
                    const int64_t e = O.off[g] + ir + R.cnt[g] * js;
                    const int64_t r = R.list[g][ir], sc = S.list[g][js];
                    Term t[1];
                    t[0] = Term{0, 0, 0, 0, false};
                    ring_of(r, sc, t);
                    if (t[0].sign != 0) {
                        const int64_t b = blk_index(P, A_T2, t[0].r, t[0].s);
                        if (b >= 0)
                            B.m0[e] = (int32_t)(((t2base + b) << 2) | (t[0].sign < 0 ? 2 : 0));
                    }
                    const int a = (int)(r % v), i = (int)(r / v);
                    const int cc = (int)(sc % v), k = (int)(sc / v);
                    const int64_t ak = blk_index(P, A_T1, a, k), ci = blk_index(P, A_T1, cc, i);
                    if (ak >= 0 && ci >= 0) {
                        B.p[0][e] = (int32_t)ak;
                        B.q[0][e] = (int32_t)ci;
                    }
                }
            });
        P.bils.push_back(std::move(B));
        RCC_ADD(P_HINTM, OK_BIL, (int)P.bils.size() - 1);
    }
    RCC_ADD(P_HINTM, OK_PROD, product(P, "HINTM_RING", 0, 2, {A_RING2, SP_VO, SP_VO},
                                      {A_cRV, SP_VO, SP_VO}, {A_H, SP_VO, SP_VO}, 1.0, 1.0, 2));
    // ---------------------------------------------------- T2EQNS
    RCC_ADD(P_T2EQNS, OK_EW, ew(P, EW_COPY, A_VVOO, A_S2));
    RCC_ADD(P_T2EQNS, OK_PROD, product(P, "T2EQNS_AZ", 0, 0, {A_TAU, SP_VV, SP_OO},
                                       {A_AZ, SP_OO, SP_OO}, {A_S2, SP_VV, SP_OO}, 1.0, 1.0));
    if (o > 1) {
        RCC_ADD(P_T2EQNS, OK_PROD, product(P, "T2EQNS_YPa", 0, 0, {A_cVOVVt, SP_VVO, SP_V},
                                           {A_T1, SP_V, SP_O}, {A_YPa, SP_VVO, SP_O}, 1.0, 0.0));
        RCC_ADD(P_T2EQNS, OK_GATHER, expand_e2(A_T2));
        RCC_ADD(P_T2EQNS, OK_PROD, product(P, "T2EQNS_YPb", 0, 0, {A_E2, SP_VVO, SP_O},
                                           {A_GO, SP_O, SP_O}, {A_YPb, SP_VVO, SP_O}, 1.0, 0.0));
        // S2[AB, {x > y}] += Y[AB, x, y] - Y[AB, y, x],
        // Y[AB, p, q] = YPa[(AB, q), p] + YPb[(AB, p), q]
        RCC_ADD(P_T2EQNS, OK_GATHER, gather(P, A_S2, 4, {A_YPa, A_YPb, A_YPa, A_YPb}, 1.0, 1.0,
            [&](int64_t ab, int64_t xy, Term* t) {
                const int x = c.oox[xy], y = c.ooy[xy];
                t[0].r = ab + c.nvvt * y; t[0].s = x; t[0].sign = 1;
                t[1].r = ab + c.nvvt * x; t[1].s = y; t[1].sign = 1;
                t[2].r = ab + c.nvvt * x; t[2].s = y; t[2].sign = -1;
                t[3].r = ab + c.nvvt * y; t[3].s = x; t[3].sign = -1;
            }));
    }
    RCC_ADD(P_T2EQNS, OK_EW, ew(P, EW_AXPBY, A_VOOO, A_WC, -1.0, 0.0));
    if (o > 1) {
        RCC_ADD(P_T2EQNS, OK_PROD, product(P, "T2EQNS_Y1", 0, 0, {A_cVOVOt, SP_VOOO, SP_V},
                                           {A_T1, SP_V, SP_O}, {A_Y1, SP_VOOO, SP_O}, 1.0, 0.0));
        // WC[bk, {x > y}] += -(Y1[(bk, y), x] - Y1[(bk, x), y])
        RCC_ADD(P_T2EQNS, OK_GATHER, gather(P, A_WC, 2, {A_Y1, A_Y1}, -1.0, 1.0,
            [&](int64_t bk, int64_t xy, Term* t) {
                const int x = c.oox[xy], y = c.ooy[xy];
                const int64_t nvo2 = (int64_t)v * o;
                t[0].r = bk + nvo2 * y; t[0].s = x; t[0].sign = 1;
                t[1].r = bk + nvo2 * x; t[1].s = y; t[1].sign = -1;
            }));
    }
    RCC_ADD(P_T2EQNS, OK_GATHER, expand_f1(A_T2));
    RCC_ADD(P_T2EQNS, OK_PROD, product(P, "T2EQNS_GV", 0, 0, {A_GV, SP_V, SP_V},
                                       {A_F1, SP_V, SP_VOO2}, {A_ZX, SP_V, SP_VOO2}, 1.0, 0.0));
    // WCt[k, (b, kl)] = WC[(b, k), kl]
    RCC_ADD(P_T2EQNS, OK_GATHER, gather(P, A_WCt, 1, {A_WC}, 1.0, 0.0,
        [&](int64_t k, int64_t sc, Term* t) {
            const int b = (int)(sc % v);
            const int64_t kl = sc / v;
            t[0].r = b + (int64_t)v * k;
            t[0].s = kl;
            t[0].sign = 1;
        }));
    RCC_ADD(P_T2EQNS, OK_PROD, product(P, "T2EQNS_WC", 0, 0, {A_T1, SP_V, SP_O},
                                       {A_WCt, SP_O, SP_VOO2}, {A_ZX, SP_V, SP_VOO2}, -1.0, 1.0));
    // S2[{a > b}, col] += ZX[a, (b, col)] - ZX[b, (a, col)]
    RCC_ADD(P_T2EQNS, OK_GATHER, gather(P, A_S2, 2, {A_ZX, A_ZX}, 1.0, 1.0,
        [&](int64_t ab, int64_t col, Term* t) {
            const int a = c.vvx[ab], b = c.vvy[ab];
            t[0].r = a; t[0].s = b + (int64_t)v * col; t[0].sign = 1;
            t[1].r = b; t[1].s = a + (int64_t)v * col; t[1].sign = -1;
        }));
    RCC_ADD(P_T2EQNS, OK_PROD, product(P, "T2EQNS_RING", 0, 0, {A_H, SP_VO, SP_VO},
                                       {A_RING, SP_VO, SP_VO}, {A_RING2, SP_VO, SP_VO}, 1.0, 0.0,
                                       2));
    // S2[ab, ij] += Q[ai,bj] - Q[bi,aj] - Q[aj,bi] + Q[bj,ai]
    RCC_ADD(P_T2EQNS, OK_GATHER, gather(P, A_S2, 4, {A_RING2, A_RING2, A_RING2, A_RING2}, 1.0,
        1.0, [&](int64_t ab, int64_t ij, Term* t) {
            const int a = c.vvx[ab], b = c.vvy[ab], i = c.oox[ij], j = c.ooy[ij];
            const int64_t ai = a + (int64_t)v * i, bj = b + (int64_t)v * j;
            const int64_t bi = b + (int64_t)v * i, aj = a + (int64_t)v * j;
            t[0].r = ai; t[0].s = bj; t[0].sign = 1;
            t[1].r = bi; t[1].s = aj; t[1].sign = -1;
            t[2].r = aj; t[2].s = bi; t[2].sign = -1;
            t[3].r = bj; t[3].s = ai; t[3].sign = 1;
        }));
    // ---------------------------------------------------- BINTM
    RCC_ADD(P_BINTM, OK_PROD, product(P, "LADDER", 0, 0, {A_VVVV, SP_VV, SP_VV},
                                      {A_TAU, SP_VV, SP_OO}, {A_S2, SP_VV, SP_OO}, 1.0, 1.0, 1));
    // Xt[(a, IJ), k] = X[(a, k), IJ]
    RCC_ADD(P_BINTM, OK_GATHER, gather(P, A_Xt, 1, {A_X}, 1.0, 0.0,
        [&](int64_t r, int64_t k, Term* t) {
            const int a = (int)(r % v);
            const int64_t ij = r / v;
            t[0].r = a + (int64_t)v * k;
            t[0].s = ij;
            t[0].sign = 1;
        }));
    RCC_ADD(P_BINTM, OK_PROD, product(P, "BINTM_X", 0, 1, {A_Xt, SP_VOO2, SP_O},
                                      {A_T1, SP_V, SP_O}, {A_ZX2, SP_VOO2, SP_V}, -1.0, 0.0));
    // S2[{a > b}, col] += ZX2[(a, col), b] - ZX2[(b, col), a]
    RCC_ADD(P_BINTM, OK_GATHER, gather(P, A_S2, 2, {A_ZX2, A_ZX2}, 1.0, 1.0,
        [&](int64_t ab, int64_t col, Term* t) {
            const int a = c.vvx[ab], b = c.vvy[ab];
            t[0].r = a + (int64_t)v * col; t[0].s = b; t[0].sign = 1;
            t[1].r = b + (int64_t)v * col; t[1].s = a; t[1].sign = -1;
        }));
    // ---------------------------------------------------- DENOM, ZCORE, MP2
    // DENOM + ZCORE of the iteration, DENOM alone, ZCORE alone, and MP2EQN's
    P.denoms.push_back(DenomOp{A_S1, A_S2, true, true});
    RCC_ADD(P_DENOM, OK_DENOM, 0);
    P.denoms.push_back(DenomOp{A_S1, A_S2, true, false});
    RCC_ADD(P_DIVIDE, OK_DENOM, 1);
    P.denoms.push_back(DenomOp{A_S1, A_S2, false, true});
    RCC_ADD(P_ZCORE, OK_DENOM, 2);
    P.denoms.push_back(DenomOp{A_T1, A_T2, true, true});
    RCC_ADD(P_MP2, OK_EW, ew(P, EW_COPY, A_FVO, A_T1));
    RCC_ADD(P_MP2, OK_EW, ew(P, EW_COPY, A_VVOO, A_T2));
    RCC_ADD(P_MP2, OK_DENOM, 3);
#undef RCC_ADD
    return 0;
}

// Denominators and frozen masks of the blocked S1 / S2 (DENOM cceqns.F:2355,
// ZCORE cceqns.F:1864): the same sums of orbital energies as the C1 kernel.
inline void build_denominators(const Plan& P, const double* eps, std::vector<double>& D1,
                               std::vector<double>& D2, std::vector<double>& F1m,
                               std::vector<double>& F2m) {
    const Ctx& c = P.c;
    const int o = c.o;
    const Arr& a1 = P.arr[A_D1];
    const Arr& a2 = P.arr[A_D2];
    D1.assign(a1.size, 0.0);
    D2.assign(a2.size, 0.0);
    F1m.assign(a1.size, 0.0);
    F2m.assign(a2.size, 0.0);
    const int N = 2 * c.G.n;
    auto fr = [&](int i) { return i < (int)c.s.frozen.size() && c.s.frozen[i] != 0; };
    for (int g = 0; g < N; ++g) {
        const Space& R1 = P.sp[a1.R];
        const Space& S1 = P.sp[a1.S];
        for (int64_t js = 0; js < S1.cnt[g]; ++js)
            for (int64_t ir = 0; ir < R1.cnt[g]; ++ir) {
                const int a = R1.list[g][ir], i = S1.list[g][js];
                const int64_t e = a1.off[g] + ir + R1.cnt[g] * js;
                D1[e] = eps[i] - eps[o + a];
                F1m[e] = fr(i) ? 1.0 : 0.0;
            }
        const Space& R2 = P.sp[a2.R];
        const Space& S2 = P.sp[a2.S];
        for (int64_t js = 0; js < S2.cnt[g]; ++js)
            for (int64_t ir = 0; ir < R2.cnt[g]; ++ir) {
                const int64_t ab = R2.list[g][ir], ij = S2.list[g][js];
                const int a = c.vvx[ab], b = c.vvy[ab], i = c.oox[ij], j = c.ooy[ij];
                const int64_t e = a2.off[g] + ir + R2.cnt[g] * js;
                D2[e] = eps[i] + eps[j] - eps[o + a] - eps[o + b];
                F2m[e] = (fr(i) || fr(j)) ? 1.0 : 0.0;
            }
    }
}

}  // namespace blk
}  // namespace rcc
