// Element-wise, permutation and reduction kernels of the relccsd HIP path.
//
// These are the HBM-bound parts of one CCSD iteration: tau (GETTAU
// ccintm.F:833), the antisymmetrised pair expansions / contractions that
// relccsd performs with its SRT* sorts (ccsort.F), the denominators
// (DENOM cceqns.F:2355) and the deterministic reductions behind DIIS
// (DIISX cceqns.F:1985) and the energy (CCENG cceqns.F:2180).
//
// Packed pair p of an n-set is (x, y) with x > y, enumerated y outer / x
// inner; the host uploads the decode table, the encode is closed form.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>
#include "rcc_internal.h"

namespace rcc {

constexpr int TPB = 256;

static inline int nblocks(int64_t n, int64_t cap = 1 << 20) {
    int64_t b = (n + TPB - 1) / TPB;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, cap));
}


__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
    return make_double2(a.x + b.x, a.y + b.y);
}

__device__ __forceinline__ double2 csub(double2 a, double2 b) {
    return make_double2(a.x - b.x, a.y - b.y);
}

__device__ __forceinline__ double2 cscale(double s, double2 a) {
    return make_double2(s * a.x, s * a.y);
}

#define GRID_LOOP(i, n) \
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); \
         i += (int64_t)gridDim.x * blockDim.x)

// 2-D element loops over a column-major array: columns over blockIdx.y,
// rows (contiguous in memory) over blockIdx.x.  The row index is a 32-bit
// value and a column's own indices are decoded once per column, so the
// bodies do no 64-bit division per element.
#define COL_LOOP(c, ncol) for (int64_t c = blockIdx.y; c < (ncol); c += gridDim.y)
#define ROW_LOOP(r, nrow)                                                     \
    for (int r = (int)(blockIdx.x * blockDim.x + threadIdx.x); r < (int)(nrow); \
         r += (int)(gridDim.x * blockDim.x))

// y-grid of a 2-D launch: the columns, at most 2048 (the loop strides)
static inline int col_grid(int64_t ncol) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(ncol, 2048));
}

// ---------------------------------------------------------------- launching
// Every element-wise kernel of the term sequence is a parameter struct P and
// a device body(const P&) with a grid-stride loop over blockIdx.x; ew_one
// runs one instance, ew_many one instance per blockIdx.z (a batch of
// handles, parameters in device memory).  With a recording active
// (g_record) the wrappers append the launch instead of issuing it, so that a
// handle's whole term sequence can be replayed for many handles at once.
thread_local std::vector<Rec>* g_record = nullptr;

template <class P>
__global__ void __launch_bounds__(TPB) ew_one(P p) { body(p); }

template <class P>
__global__ void __launch_bounds__(TPB) ew_many(const P* __restrict__ ps) { body(ps[blockIdx.z]); }

template <class P>
static void ew_launch(int kind, const P& p, int64_t n, int gy, hipStream_t st) {
    if (g_record) {
        Rec r;
        r.kind = kind;
        r.n = n;
        r.gy = gy;
        r.blob.assign((const unsigned char*)&p, (const unsigned char*)&p + sizeof(P));
        g_record->push_back(std::move(r));
        return;
    }
    if (n <= 0 || gy <= 0) return;
    hipLaunchKernelGGL(ew_one<P>, dim3(nblocks(n), gy), dim3(TPB), 0, st, p);
}

// ---------------------------------------------------------------- tau
struct GettauP {
    const double2* t1; const double2* t2; double2* tau; double s2;
    const int2* vv; int64_t nvvt; const int2* oo; int64_t noot; int nv;
};
__device__ __forceinline__ void body(const GettauP& p) {
    COL_LOOP(ij, p.noot) ROW_LOOP(ab, p.nvvt) {
        const int64_t e = ab + p.nvvt * ij;
        const int2 q1 = p.vv[ab], q2 = p.oo[ij];
        const int a = q1.x, b = q1.y, i = q2.x, j = q2.y, nv = p.nv;
        double2 v = cscale(p.s2, p.t2[e]);
        v = cadd(v, cmul(p.t1[a + nv * i], p.t1[b + nv * j]));
        v = csub(v, cmul(p.t1[a + nv * j], p.t1[b + nv * i]));
        p.tau[e] = v;
    }
}

void k_gettau(const double2* t1, const double2* t2, double2* tau, double s2, PairTab vv,
              PairTab oo, int nv, hipStream_t st) {
    const GettauP p{t1, t2, tau, s2, vv.tab, vv.npair, oo.tab, oo.npair, nv};
    ew_launch(RK_GETTAU, p, vv.npair, col_grid(oo.npair), st);
}

// ---------------------------------------------------------------- expansions
struct ExpandE2P { const double2* P; double2* out; int64_t R; int no; };
__device__ __forceinline__ void body(const ExpandE2P& p) {
    const int no = p.no;
    COL_LOOP(q, (int64_t)no * no) {
        const int y = (int)(q % no), x = (int)(q / no);
        ROW_LOOP(r, p.R) {
            double2 v = make_double2(0.0, 0.0);
            if (x > y) v = p.P[r + p.R * pair_of(x, y, no)];
            else if (x < y) { v = p.P[r + p.R * pair_of(y, x, no)]; v.x = -v.x; v.y = -v.y; }
            p.out[r + p.R * q] = v;
        }
    }
}

void k_expand_e2(const double2* P, double2* out, int64_t R, PairTab oo, int no, hipStream_t st) {
    (void)oo;
    ew_launch(RK_EXPAND_E2, ExpandE2P{P, out, R, no}, R, col_grid((int64_t)no * no), st);
}

// out[(a,d), col] = +-P(ad, col) (0 on a = d).  Tiled like the ring form: a
// 32 x 32 (a, d) tile of one column, the a > d half read along a and the
// a < d half along d into LDS, the tile written along a.
constexpr int F1_T = 32;
struct ExpandF1P { const double2* P; double2* out; int64_t ncol; int nv; int64_t nvvt; };
__device__ __forceinline__ void body(const ExpandF1P& p) {
    __shared__ double2 tile[F1_T][F1_T + 1];
    const int nv = p.nv;
    const int64_t nv2 = (int64_t)nv * nv;
    const int nt = (nv + F1_T - 1) / F1_T;
    const int tx = (int)threadIdx.x % F1_T, ty = (int)threadIdx.x / F1_T;
    constexpr int TY = TPB / F1_T;
    for (int64_t col = blockIdx.y; col < p.ncol; col += gridDim.y) {
        const double2* src = p.P + p.nvvt * col;
        for (int t = blockIdx.x; t < nt * nt; t += gridDim.x) {
            const int a0 = (t % nt) * F1_T, d0 = (t / nt) * F1_T;
            for (int y = ty; y < F1_T; y += TY) {
                int a = a0 + tx, d = d0 + y;            // a > d: contiguous in a
                if (a < nv && a > d) tile[y][tx] = src[pair_of(a, d, nv)];
                a = a0 + y;
                d = d0 + tx;                            // a < d: contiguous in d
                if (d < nv && a < d) tile[tx][y] = src[pair_of(d, a, nv)];
            }
            __syncthreads();
            for (int y = ty; y < F1_T; y += TY) {
                const int a = a0 + tx, d = d0 + y;
                if (a >= nv || d >= nv) continue;
                double2 v = make_double2(0.0, 0.0);
                if (a > d) v = tile[y][tx];
                else if (a < d) { v = tile[y][tx]; v.x = -v.x; v.y = -v.y; }
                p.out[(a + (int64_t)nv * d) + nv2 * col] = v;
            }
            __syncthreads();
        }
    }
}

void k_expand_f1(const double2* P, double2* out, int64_t ncol, PairTab vv, int nv,
                 hipStream_t st) {
    const int64_t nt = (nv + F1_T - 1) / F1_T;
    ew_launch(RK_EXPAND_F1, ExpandF1P{P, out, ncol, nv, vv.npair}, nt * nt * TPB,
              col_grid(ncol), st);
}

struct ExpandRingP {
    const double2* T2; const double2* t1; double s; double2* out; int nv; int no; int64_t nvvt;
};
// out[(a,i),(c,k)] = +-s T2(ac, ik) - t1(a,k) t1(c,i), column-major.  The
// workgroup takes a 32 x 32 tile of (a, c) at one (i, k).  In the packed
// y-major triangle pair_of(a, c) runs contiguously in a where a > c and in c
// where a < c, so the tile's T2 elements are read in those two coalesced
// halves into LDS (as [c][a]) and written out along a.
constexpr int RING_T = 32;
__device__ __forceinline__ void body(const ExpandRingP& p) {
    __shared__ double2 tile[RING_T][RING_T + 1];
    const int nv = p.nv, no = p.no;
    const int64_t nvo = (int64_t)nv * no;
    const int nt = (nv + RING_T - 1) / RING_T;
    const int ntiles = nt * no;     // (a tile, i) row tiles; (c tile, k) column tiles alike
    const int tx = (int)threadIdx.x % RING_T, ty = (int)threadIdx.x / RING_T;
    constexpr int TY = TPB / RING_T;
    for (int ct = blockIdx.y; ct < ntiles; ct += gridDim.y) {
        const int c0 = (ct % nt) * RING_T, k = ct / nt;
        for (int rt = blockIdx.x; rt < ntiles; rt += gridDim.x) {
            const int a0 = (rt % nt) * RING_T, i = rt / nt;
            if (i != k) {
                const double2* src =
                    p.T2 + p.nvvt * (i > k ? pair_of(i, k, no) : pair_of(k, i, no));
                for (int y = ty; y < RING_T; y += TY) {
                    int a = a0 + tx, c = c0 + y;        // a > c: contiguous in a
                    if (a < nv && a > c) tile[y][tx] = src[pair_of(a, c, nv)];
                    a = a0 + y;
                    c = c0 + tx;                        // a < c: contiguous in c
                    if (c < nv && a < c) tile[tx][y] = src[pair_of(c, a, nv)];
                }
                __syncthreads();
            }
            for (int y = ty; y < RING_T; y += TY) {
                const int a = a0 + tx, c = c0 + y;
                if (a >= nv || c >= nv) continue;
                double2 v = make_double2(0.0, 0.0);
                if (a != c && i != k) {
                    const double sg = ((a > c) == (i > k)) ? p.s : -p.s;
                    v = cscale(sg, tile[y][tx]);
                }
                if (p.t1) v = csub(v, cmul(p.t1[a + nv * k], p.t1[c + nv * i]));
                p.out[(a + (int64_t)nv * i) + nvo * (c + (int64_t)nv * k)] = v;
            }
            __syncthreads();
        }
    }
}

void k_expand_ring(const double2* T2, const double2* t1, double s, double2* out, PairTab vv,
                   PairTab oo, int nv, int no, hipStream_t st) {
    (void)oo;
    const int64_t ntiles = (int64_t)((nv + RING_T - 1) / RING_T) * no;
    ew_launch(RK_EXPAND_RING, ExpandRingP{T2, t1, s, out, nv, no, vv.npair}, ntiles * TPB,
              col_grid(ntiles), st);
}

__global__ void expand_cols_kernel(const double2* __restrict__ P, double2* __restrict__ out,
                                   int64_t R, int nv) {
    const int64_t nv2 = (int64_t)nv * nv;
    GRID_LOOP(o, R * nv2) {
        const int64_t row = o % R;
        const int64_t q = o / R;
        const int c = (int)(q % nv), d = (int)(q / nv);
        double2 v = make_double2(0.0, 0.0);
        if (c > d) v = P[row + R * pair_of(c, d, nv)];
        else if (c < d) { v = P[row + R * pair_of(d, c, nv)]; v.x = -v.x; v.y = -v.y; }
        out[o] = v;
    }
}

void k_expand_cols(const double2* P, double2* out, int64_t R, PairTab vv, int nv,
                   hipStream_t st) {
    (void)vv;
    const int64_t n = R * nv * nv;
    if (n == 0) return;
    hipLaunchKernelGGL(expand_cols_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, P, out, R, nv);
}

// ---------------------------------------------------------------- antisymmetrising packs
struct PackPijP { const double2* Y; double2* S; int64_t NR; double alpha; const int2* oo;
                  int64_t noot; int no; };
__device__ __forceinline__ void body(const PackPijP& p) {
    COL_LOOP(ij, p.noot) ROW_LOOP(r, p.NR) {
        const int64_t e = r + p.NR * ij;
        const int2 q = p.oo[ij];
        const double2 y1 = p.Y[r + p.NR * (q.x + (int64_t)p.no * q.y)];
        const double2 y2 = p.Y[r + p.NR * (q.y + (int64_t)p.no * q.x)];
        double2 s = p.S[e];
        s.x += p.alpha * (y1.x - y2.x);
        s.y += p.alpha * (y1.y - y2.y);
        p.S[e] = s;
    }
}

void k_pack_pij(const double2* Y, double2* S, int64_t NR, double alpha, PairTab oo, int no,
                hipStream_t st) {
    ew_launch(RK_PACK_PIJ, PackPijP{Y, S, NR, alpha, oo.tab, oo.npair, no}, NR,
              col_grid(oo.npair), st);
}

struct PackPabP { const double2* Y; double2* S; int64_t ncol; double alpha; const int2* vv;
                  int64_t nvvt; int nv; };
// The packs walk 32 x 32 (a, b) tiles with a-tile >= b-tile (tile t of the
// triangle); the operands contiguous in b are read along b into LDS and the
// packed pairs, contiguous in a, are updated along a.
constexpr int PK_T = 32;
__device__ __forceinline__ void tri_tile(int t, int& ta, int& tb) {
    ta = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while (ta * (ta + 1) / 2 > t) --ta;
    while ((ta + 1) * (ta + 2) / 2 <= t) ++ta;
    tb = t - ta * (ta + 1) / 2;
}
static inline int64_t pack_tiles(int nv) {
    const int64_t nt = (nv + PK_T - 1) / PK_T;
    return nt * (nt + 1) / 2;
}

__device__ __forceinline__ void body(const PackPabP& p) {
    __shared__ double2 tile[PK_T][PK_T + 1];
    const int nv = p.nv;
    const int64_t nv2 = (int64_t)nv * nv;
    const int nt = (nv + PK_T - 1) / PK_T;
    const int tx = (int)threadIdx.x % PK_T, ty = (int)threadIdx.x / PK_T;
    constexpr int TY = TPB / PK_T;
    for (int64_t col = blockIdx.y; col < p.ncol; col += gridDim.y) {
        const double2* Y = p.Y + nv2 * col;
        for (int t = blockIdx.x; t < nt * (nt + 1) / 2; t += gridDim.x) {
            int ta, tb;
            tri_tile(t, ta, tb);
            const int a0 = ta * PK_T, b0 = tb * PK_T;
            for (int y = ty; y < PK_T; y += TY) {
                const int a = a0 + y, b = b0 + tx;      // Y(b, a): contiguous in b
                if (a < nv && a > b) tile[y][tx] = Y[b + (int64_t)nv * a];
            }
            __syncthreads();
            for (int y = ty; y < PK_T; y += TY) {
                const int a = a0 + tx, b = b0 + y;
                if (a >= nv || a <= b) continue;
                const int64_t e = pair_of(a, b, nv) + p.nvvt * col;
                const double2 y1 = Y[a + (int64_t)nv * b];
                const double2 y2 = tile[tx][y];
                double2 s = p.S[e];
                s.x += p.alpha * (y1.x - y2.x);
                s.y += p.alpha * (y1.y - y2.y);
                p.S[e] = s;
            }
            __syncthreads();
        }
    }
}

void k_pack_pab(const double2* Y, double2* S, int64_t ncol, double alpha, PairTab vv, int nv,
                hipStream_t st) {
    ew_launch(RK_PACK_PAB, PackPabP{Y, S, ncol, alpha, vv.tab, vv.npair, nv},
              pack_tiles(nv) * TPB, col_grid(ncol), st);
}

struct PackPabijP { const double2* Q; double2* S; const int2* vv; int64_t nvvt; const int2* oo;
                    int64_t noot; int nv; int no; double alpha; };
__device__ __forceinline__ void body(const PackPabijP& p) {
    __shared__ double2 t2[PK_T][PK_T + 1], t4[PK_T][PK_T + 1];
    const int nv = p.nv;
    const int64_t nvo = (int64_t)nv * p.no;
    const int nt = (nv + PK_T - 1) / PK_T;
    const int tx = (int)threadIdx.x % PK_T, ty = (int)threadIdx.x / PK_T;
    constexpr int TY = TPB / PK_T;
    for (int64_t ij = blockIdx.y; ij < p.noot; ij += gridDim.y) {
        const int2 q2 = p.oo[ij];
        const int64_t ci = (int64_t)nv * q2.x, cj = (int64_t)nv * q2.y;
        for (int t = blockIdx.x; t < nt * (nt + 1) / 2; t += gridDim.x) {
            int ta, tb;
            tri_tile(t, ta, tb);
            const int a0 = ta * PK_T, b0 = tb * PK_T;
            for (int y = ty; y < PK_T; y += TY) {
                const int a = a0 + y, b = b0 + tx;      // Q((b,i),(a,j)), Q((b,j),(a,i))
                if (a < nv && a > b) {
                    t2[y][tx] = p.Q[(b + ci) + nvo * (a + cj)];
                    t4[y][tx] = p.Q[(b + cj) + nvo * (a + ci)];
                }
            }
            __syncthreads();
            for (int y = ty; y < PK_T; y += TY) {
                const int a = a0 + tx, b = b0 + y;
                if (a >= nv || a <= b) continue;
                const int64_t e = pair_of(a, b, nv) + p.nvvt * ij;
                const double2 v1 = p.Q[(a + ci) + nvo * (b + cj)], v2 = t2[tx][y];
                const double2 v3 = p.Q[(a + cj) + nvo * (b + ci)], v4 = t4[tx][y];
                double2 s = p.S[e];
                s.x += p.alpha * (v1.x - v2.x - v3.x + v4.x);
                s.y += p.alpha * (v1.y - v2.y - v3.y + v4.y);
                p.S[e] = s;
            }
            __syncthreads();
        }
    }
}

void k_pack_pabij(const double2* Q, double2* S, PairTab vv, PairTab oo, int nv, int no,
                  hipStream_t st, double alpha) {
    ew_launch(RK_PACK_PABIJ, PackPabijP{Q, S, vv.tab, vv.npair, oo.tab, oo.npair, nv, no, alpha},
              pack_tiles(nv) * TPB, col_grid(oo.npair), st);
}

// ---------------------------------------------------------------- ring permutation
struct PermAkciP { const double2* X; double2* H; int nv; int no; int accumulate; };
__device__ __forceinline__ void body(const PermAkciP& p) {
    const int nv = p.nv, no = p.no;
    const int64_t nvo = (int64_t)nv * no;
    COL_LOOP(cc, nvo) ROW_LOOP(r, nvo) {
        const int c = (int)(cc % nv), k = (int)(cc / nv);
        const int a = r % nv, i = r / nv;
        const int64_t o = r + nvo * cc;
        const double2 x = p.X[(a + (int64_t)nv * k) + nvo * (c + (int64_t)nv * i)];
        p.H[o] = p.accumulate ? cadd(p.H[o], x) : x;
    }
}

void k_perm_akci(const double2* X, double2* H, int nv, int no, bool accumulate,
                 hipStream_t st) {
    const int64_t nvo = (int64_t)nv * no;
    ew_launch(RK_PERM_AKCI, PermAkciP{X, H, nv, no, accumulate ? 1 : 0}, nvo, col_grid(nvo), st);
}

// ---------------------------------------------------------------- traces
struct TraceT1P { const double2* X; double2* S1; int nv; int no; };
__device__ __forceinline__ void body(const TraceT1P& p) {
    const int nv = p.nv, no = p.no;
    const int64_t nvo = (int64_t)nv * no;
    GRID_LOOP(e, nvo) {
        const int a = (int)(e % nv), i = (int)(e / nv);
        double2 s = p.S1[e];
        for (int k = 0; k < no; ++k) {
            if (k == i) continue;
            const double2 x =
                p.X[(a + (int64_t)nv * k) + nvo * (i > k ? pair_of(i, k, no) : pair_of(k, i, no))];
            if (i > k) { s.x += x.x; s.y += x.y; }
            else { s.x -= x.x; s.y -= x.y; }
        }
        p.S1[e] = s;
    }
}

void k_trace_t1eqn2(const double2* X, double2* S1, PairTab oo, int nv, int no, hipStream_t st) {
    (void)oo;
    ew_launch(RK_TRACE_T1, TraceT1P{X, S1, nv, no}, (int64_t)nv * no, 1, st);
}

struct TraceGvP { const double2* H1; double2* GV; double alpha; int nv; int no; };
__device__ __forceinline__ void body(const TraceGvP& p) {
    const int nv = p.nv, no = p.no;
    const int64_t nvo = (int64_t)nv * no;
    GRID_LOOP(e, (int64_t)nv * nv) {
        const int a = (int)(e % nv), c = (int)(e / nv);
        double sr = 0.0, si = 0.0;
        for (int k = 0; k < no; ++k) {
            const double2 x = p.H1[(a + (int64_t)nv * k) + nvo * (c + (int64_t)nv * k)];
            sr += x.x;
            si += x.y;
        }
        double2 g = p.GV[e];
        g.x += p.alpha * sr;
        g.y += p.alpha * si;
        p.GV[e] = g;
    }
}

void k_trace_gv(const double2* H1, double2* GV, double alpha, int nv, int no, hipStream_t st) {
    ew_launch(RK_TRACE_GV, TraceGvP{H1, GV, alpha, nv, no}, (int64_t)nv * nv, 1, st);
}

// ---------------------------------------------------------------- adjoints of the traces
// Transpose of trace_t1eqn2: Xb[(a,y), p] += S1b[a,x], Xb[(a,x), p] -= S1b[a,y]
// for p = (x > y).  Each X element is written by one thread.
__global__ void trace_t1eqn2_adj_kernel(const double2* __restrict__ S1b, double2* __restrict__ Xb,
                                        const int2* __restrict__ oo, int64_t noot, int nv,
                                        int no) {
    const int64_t nvo = (int64_t)nv * no;
    GRID_LOOP(e, (int64_t)nv * noot) {
        const int a = (int)(e % nv);
        const int64_t p = e / nv;
        const int2 q = oo[p];
        double2* xy = Xb + (a + (int64_t)nv * q.y) + nvo * p;
        double2* xx = Xb + (a + (int64_t)nv * q.x) + nvo * p;
        *xy = cadd(*xy, S1b[a + (int64_t)nv * q.x]);
        *xx = csub(*xx, S1b[a + (int64_t)nv * q.y]);
    }
}

void k_trace_t1eqn2_adj(const double2* S1b, double2* Xb, PairTab oo, int nv, int no,
                        hipStream_t st) {
    const int64_t n = (int64_t)nv * oo.npair;
    if (n == 0) return;
    hipLaunchKernelGGL(trace_t1eqn2_adj_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, S1b, Xb,
                       oo.tab, oo.npair, nv, no);
}

// Transpose of trace_gv: H1b[(a,k),(c,k)] += alpha * GVb[a,c] for every k.
__global__ void trace_gv_adj_kernel(const double2* __restrict__ GVb, double2* __restrict__ H1b,
                                    double alpha, int nv, int no) {
    const int64_t nvo = (int64_t)nv * no;
    const int64_t nv2 = (int64_t)nv * nv;
    GRID_LOOP(e, nv2 * no) {
        const int a = (int)(e % nv);
        const int c = (int)((e / nv) % nv);
        const int k = (int)(e / nv2);
        double2* h = H1b + (a + (int64_t)nv * k) + nvo * (c + (int64_t)nv * k);
        const double2 g = GVb[a + (int64_t)nv * c];
        h->x += alpha * g.x;
        h->y += alpha * g.y;
    }
}

void k_trace_gv_adj(const double2* GVb, double2* H1b, double alpha, int nv, int no,
                    hipStream_t st) {
    const int64_t n = (int64_t)nv * nv * no;
    if (n == 0) return;
    hipLaunchKernelGGL(trace_gv_adj_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, GVb, H1b, alpha,
                       nv, no);
}

__global__ void hadamard_kernel(int64_t n, const double2* __restrict__ x,
                                const double2* __restrict__ y, double2* __restrict__ z) {
    GRID_LOOP(i, n) z[i] = cmul(x[i], y[i]);
}

void k_hadamard(int64_t n, const double2* x, const double2* y, double2* z, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(hadamard_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, n, x, y, z);
}

// ---------------------------------------------------------------- denominators
struct Denom1P { double2* S1; const double* eps; const int* frozen; int nv; int no; };
__device__ __forceinline__ void body(const Denom1P& p) {
    GRID_LOOP(e, (int64_t)p.nv * p.no) {
        const int a = (int)(e % p.nv), i = (int)(e / p.nv);
        double2 s = p.S1[e];
        if (p.eps) {
            const double d = p.eps[i] - p.eps[p.no + a];
            s.x /= d;
            s.y /= d;
        }
        if (p.frozen && p.frozen[i]) s = make_double2(0.0, 0.0);
        p.S1[e] = s;
    }
}

struct Denom2P { double2* S2; const double* eps; const int* frozen; const int2* vv; int64_t nvvt;
                 const int2* oo; int64_t noot; int no; };
__device__ __forceinline__ void body(const Denom2P& p) {
    COL_LOOP(ij, p.noot) ROW_LOOP(ab, p.nvvt) {
        const int64_t e = ab + p.nvvt * ij;
        const int2 q1 = p.vv[ab], q2 = p.oo[ij];
        double2 s = p.S2[e];
        if (p.eps) {
            const double d = p.eps[q2.x] + p.eps[q2.y] - p.eps[p.no + q1.x] - p.eps[p.no + q1.y];
            s.x /= d;
            s.y /= d;
        }
        if (p.frozen && (p.frozen[q2.x] || p.frozen[q2.y])) s = make_double2(0.0, 0.0);
        p.S2[e] = s;
    }
}

void k_denom(double2* S1, double2* S2, const double* eps, const int* frozen, PairTab vv,
             PairTab oo, int nv, int no, hipStream_t st) {
    if (S1) ew_launch(RK_DENOM1, Denom1P{S1, eps, frozen, nv, no}, (int64_t)nv * no, 1, st);
    if (S2)
        ew_launch(RK_DENOM2, Denom2P{S2, eps, frozen, vv.tab, vv.npair, oo.tab, oo.npair, no},
                  vv.npair, col_grid(oo.npair), st);
}

// ---------------------------------------------------------------- small BLAS-1
struct AxpbyP { int64_t n; double2 a; const double2* x; double2 b; double2* y; };
__device__ __forceinline__ void body(const AxpbyP& p) {
    GRID_LOOP(e, p.n) {
        double2 r = make_double2(0.0, 0.0);
        if (p.b.x != 0.0 || p.b.y != 0.0) r = cmul(p.b, p.y[e]);
        if (p.x) r = cadd(r, cmul(p.a, p.x[e]));
        p.y[e] = r;
    }
}

void k_axpby(int64_t n, double2 a, const double2* x, double2 b, double2* y, hipStream_t st) {
    ew_launch(RK_AXPBY, AxpbyP{n, a, x, b, y}, n, 1, st);
}

struct ConjP { int64_t n; const double2* x; double2* y; };
__device__ __forceinline__ void body(const ConjP& p) {
    GRID_LOOP(e, p.n) { p.y[e] = make_double2(p.x[e].x, -p.x[e].y); }
}

void k_conj(int64_t n, const double2* x, double2* y, hipStream_t st) {
    ew_launch(RK_CONJ, ConjP{n, x, y}, n, 1, st);
}

struct CopyP { int64_t n; const double2* x; double2* y; };
__device__ __forceinline__ void body(const CopyP& p) {
    GRID_LOOP(e, p.n) { p.y[e] = p.x[e]; }
}

void k_copy(int64_t n, const double2* x, double2* y, hipStream_t st) {
    ew_launch(RK_COPY, CopyP{n, x, y}, n, 1, st);
}

struct ZeroDiagP { double2* M; int n; };
__device__ __forceinline__ void body(const ZeroDiagP& p) {
    GRID_LOOP(e, p.n) { p.M[e + (int64_t)p.n * e] = make_double2(0.0, 0.0); }
}

void k_zero_diag(double2* M, int n, hipStream_t st) {
    ew_launch(RK_ZERO_DIAG, ZeroDiagP{M, n}, n, 1, st);
}

__global__ void gather_kernel(const double2* __restrict__ in, const int64_t* __restrict__ idx,
                              const signed char* __restrict__ sgn, int conj,
                              double2* __restrict__ out, int64_t n) {
    GRID_LOOP(e, n) {
        const int64_t s = idx[e];
        double2 v = make_double2(0.0, 0.0);
        if (s >= 0) {
            v = in[s];
            if (conj) v.y = -v.y;
            if (sgn[e] < 0) { v.x = -v.x; v.y = -v.y; }
        }
        out[e] = v;
    }
}

void k_gather(const double2* in, const int64_t* idx, const signed char* sgn, bool conj,
              double2* out, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(gather_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, in, idx, sgn,
                       conj ? 1 : 0, out, n);
}

// ---------------------------------------------------------------- reductions
constexpr int MAXV = 16;
struct PtrPack {
    const double2* p[MAXV];
};
struct CoefPack {
    const double2* p[MAXV];
    double c[MAXV];
};

// Reductions: each thread accumulates its strided elements, a wave adds its
// 64 lanes with shuffles (fixed butterfly order), the waves' sums meet in LDS
// and one lane adds them in wave order -- so the result does not depend on
// scheduling.
constexpr int NWAVE = TPB / 64;

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    return x;     // lane 0 holds the sum
}

// Block b reduces elements [b*chunk, min(n,(b+1)*chunk)) of sum conj(x_v) * y
// for every v, reading y once for all of them.
struct ZdotPartP { PtrPack xs; int nx; const double2* y; int64_t n; int64_t chunk;
                   double2* part; int nblk; };
__device__ __forceinline__ void body(const ZdotPartP& p) {
    __shared__ double sh[2][MAXV][NWAVE];
    if ((int)blockIdx.x >= p.nblk) return;      // whole workgroup (batched launches)
    const int64_t beg = blockIdx.x * p.chunk;
    const int64_t end = std::min<int64_t>(p.n, beg + p.chunk);
    double sr[MAXV], si[MAXV];
#pragma unroll
    for (int v = 0; v < MAXV; ++v) sr[v] = si[v] = 0.0;
    for (int64_t e = beg + threadIdx.x; e < end; e += TPB) {
        const double2 b = p.y[e];
#pragma unroll
        for (int v = 0; v < MAXV; ++v)
            if (v < p.nx) {
                const double2 a = p.xs.p[v][e];
                sr[v] += a.x * b.x + a.y * b.y;
                si[v] += a.x * b.y - a.y * b.x;
            }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < MAXV; ++v)
        if (v < p.nx) {
            const double r = wave_sum(sr[v]), i = wave_sum(si[v]);
            if (lane == 0) {
                sh[0][v][w] = r;
                sh[1][v][w] = i;
            }
        }
    __syncthreads();
    if ((int)threadIdx.x < p.nx) {
        double r = 0.0, i = 0.0;
#pragma unroll
        for (int q = 0; q < NWAVE; ++q) {
            r += sh[0][threadIdx.x][q];
            i += sh[1][threadIdx.x][q];
        }
        p.part[threadIdx.x * p.nblk + blockIdx.x] = make_double2(r, i);
    }
}

// One workgroup per reduced value: strided partial sums, wave shuffles, then
// the waves in order.
struct FinishP { const double2* part; int nblk; int nv; double2* dst; };
__device__ __forceinline__ void body(const FinishP& p) {
    __shared__ double sh[2][NWAVE];
    const int v = blockIdx.x;
    if (v >= p.nv) return;
    double sr = 0.0, si = 0.0;
    for (int b = threadIdx.x; b < p.nblk; b += TPB) {
        sr += p.part[v * p.nblk + b].x;
        si += p.part[v * p.nblk + b].y;
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][w] = sr;
        sh[1][w] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0, i = 0.0;
#pragma unroll
        for (int q = 0; q < NWAVE; ++q) {
            r += sh[0][q];
            i += sh[1][q];
        }
        p.dst[v] = make_double2(r, i);
    }
}

// partial-sum blocks for n elements: about 4096 elements each, at most nblk
int red_blocks(int64_t n, int nblk) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(nblk, (n + 4095) / 4096));
}

void zdot_params(const double2* const* xs, int nx, const double2* y, int64_t n, double2* dst,
                 double2* part, int nblk, void* zp, void* fp) {
    ZdotPartP z{};
    for (int v = 0; v < nx && v < MAXV; ++v) z.xs.p[v] = xs[v];
    z.nx = nx;
    z.y = y;
    z.n = n;
    z.nblk = red_blocks(n, nblk);
    z.chunk = std::max<int64_t>((n + z.nblk - 1) / z.nblk, 1);
    z.part = part;
    *(ZdotPartP*)zp = z;
    *(FinishP*)fp = FinishP{part, z.nblk, nx, dst};
}

void k_zdotc_multi(const double2* const* xs, int nx, const double2* y, int64_t n, double2* dst,
                   double2* part, int nblk, hipStream_t st) {
    if (nx <= 0) return;
    ZdotPartP z;
    FinishP f;
    zdot_params(xs, nx, y, n, dst, part, nblk, &z, &f);
    hipLaunchKernelGGL(ew_one<ZdotPartP>, dim3(z.nblk), dim3(TPB), 0, st, z);
    hipLaunchKernelGGL(ew_one<FinishP>, dim3(nx), dim3(TPB), 0, st, f);
}

// dst[0] = sum x, dst[1] = (sum |x|^2, 0)
struct SumNormPartP { const double2* x; int64_t n; int64_t chunk; double2* part; int nblk; };
__device__ __forceinline__ void body(const SumNormPartP& p) {
    __shared__ double sh[3][NWAVE];
    if ((int)blockIdx.x >= p.nblk) return;
    const int64_t beg = blockIdx.x * p.chunk;
    const int64_t end = std::min<int64_t>(p.n, beg + p.chunk);
    double sr = 0.0, si = 0.0, nn = 0.0;
    for (int64_t e = beg + threadIdx.x; e < end; e += TPB) {
        const double2 a = p.x[e];
        sr += a.x;
        si += a.y;
        nn += a.x * a.x + a.y * a.y;
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    nn = wave_sum(nn);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][w] = sr;
        sh[1][w] = si;
        sh[2][w] = nn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0, i = 0.0, q2 = 0.0;
#pragma unroll
        for (int q = 0; q < NWAVE; ++q) {
            r += sh[0][q];
            i += sh[1][q];
            q2 += sh[2][q];
        }
        p.part[blockIdx.x] = make_double2(r, i);
        p.part[p.nblk + blockIdx.x] = make_double2(q2, 0.0);
    }
}

void sumnorm_params(const double2* x, int64_t n, double2* dst, double2* part, int nblk, void* sp,
                    void* fp) {
    SumNormPartP s{x, n, 1, part, red_blocks(n, nblk)};
    s.chunk = std::max<int64_t>((n + s.nblk - 1) / s.nblk, 1);
    *(SumNormPartP*)sp = s;
    *(FinishP*)fp = FinishP{part, s.nblk, 2, dst};
}

void k_sum_and_norm(const double2* x, int64_t n, double2* dst, double2* part, int nblk,
                    hipStream_t st) {
    SumNormPartP s;
    FinishP f;
    sumnorm_params(x, n, dst, part, nblk, &s, &f);
    hipLaunchKernelGGL(ew_one<SumNormPartP>, dim3(s.nblk), dim3(TPB), 0, st, s);
    hipLaunchKernelGGL(ew_one<FinishP>, dim3(2), dim3(TPB), 0, st, f);
}

struct LincombP { CoefPack cp; int nx; double2* y; int64_t n; };
__device__ __forceinline__ void body(const LincombP& p) {
    GRID_LOOP(e, p.n) {
        double sr = 0.0, si = 0.0;
        for (int v = 0; v < p.nx; ++v) {
            const double2 a = p.cp.p[v][e];
            sr += p.cp.c[v] * a.x;
            si += p.cp.c[v] * a.y;
        }
        p.y[e] = make_double2(sr, si);
    }
}

void lincomb_params(const double2* const* xs, const double* coef, int nx, double2* y, int64_t n,
                    void* lp) {
    LincombP l{};
    for (int v = 0; v < nx && v < MAXV; ++v) {
        l.cp.p[v] = xs[v];
        l.cp.c[v] = coef[v];
    }
    l.nx = nx;
    l.y = y;
    l.n = n;
    *(LincombP*)lp = l;
}

void k_lincomb(const double2* const* xs, const double* coef, int nx, double2* y, int64_t n,
               hipStream_t st) {
    if (n <= 0) return;
    LincombP l;
    lincomb_params(xs, coef, nx, y, n, &l);
    hipLaunchKernelGGL(ew_one<LincombP>, dim3(nblocks(n)), dim3(TPB), 0, st, l);
}

// DIISX's error vector and store in one pass: d = S - T (as y = 1 S + (-1) T),
// amp = S.
struct DiisPrepP { const double2* S; const double2* T; double2* d; double2* amp; int64_t n; };
__device__ __forceinline__ void body(const DiisPrepP& p) {
    GRID_LOOP(e, p.n) {
        const double2 s = p.S[e];
        p.d[e] = cadd(cmul(make_double2(1.0, 0.0), s), cmul(make_double2(-1.0, 0.0), p.T[e]));
        p.amp[e] = s;
    }
}

void diis_prep_params(const double2* S, const double2* T, double2* d, double2* amp, int64_t n,
                      void* pp) {
    *(DiisPrepP*)pp = DiisPrepP{S, T, d, amp, n};
}

// ---------------------------------------------------------------- symmetry-blocked engine
// The element-wise work of the blocked iteration (rcc_blocked.hip): gathers
// through precomputed index maps, the bilinear gather of tau and the ring
// intermediate, denominators per blocked element and the K-split reductions
// of grouped products.  Map entries are (index << 2) | (negate << 1) | conj,
// -1 for a zero term, stored term-major.
struct MapSumP {
    double2* out; const double2* src[4]; int slot[4]; const int32_t* map; int64_t n; int T;
    double alpha, beta;
};
__device__ __forceinline__ void body(const MapSumP& p) {
    GRID_LOOP(e, p.n) {
        double sr = 0.0, si = 0.0;
        for (int t = 0; t < p.T; ++t) {
            const int32_t m = p.map[(int64_t)t * p.n + e];
            if (m < 0) continue;
            const double2 v = p.src[p.T <= 4 ? p.slot[t] : 0][m >> 2];
            const double vi = (m & 1) ? -v.y : v.y;
            if (m & 2) { sr -= v.x; si -= vi; }
            else { sr += v.x; si += vi; }
        }
        double2 o = make_double2(p.alpha * sr, p.alpha * si);
        if (p.beta != 0.0) {
            const double2 c = p.out[e];
            o.x += p.beta * c.x;
            o.y += p.beta * c.y;
        }
        p.out[e] = o;
    }
}

void k_map_sum(double2* out, const double2* const* src, const int* slot, int nsrc,
               const int32_t* map, int64_t n, int T, double alpha, double beta, hipStream_t st) {
    MapSumP p{};
    p.out = out;
    for (int i = 0; i < 4; ++i) {
        p.src[i] = i < nsrc ? src[i] : nullptr;
        p.slot[i] = i < T && T <= 4 ? slot[i] : 0;
    }
    p.map = map;
    p.n = n;
    p.T = T;
    p.alpha = alpha;
    p.beta = beta;
    ew_launch(RK_MAPSUM, p, n, 1, st);
}

struct MapBilP {
    double2* out; const double2* src0; const double2* A; const double2* B;
    const int32_t* m0; const int32_t* p0; const int32_t* q0; const int32_t* p1; const int32_t* q1;
    int64_t n; int np; double c0, sg0, sg1;
};
__device__ __forceinline__ void body(const MapBilP& p) {
    GRID_LOOP(e, p.n) {
        double2 v = make_double2(0.0, 0.0);
        const int32_t m = p.m0[e];
        if (m >= 0) {
            double2 w = p.src0[m >> 2];
            if (m & 1) w.y = -w.y;
            if (m & 2) { w.x = -w.x; w.y = -w.y; }
            v = cscale(p.c0, w);
        }
        if (p.np > 0 && p.p0[e] >= 0) {
            const double2 q = cmul(p.A[p.p0[e]], p.B[p.q0[e]]);
            v.x += p.sg0 * q.x;
            v.y += p.sg0 * q.y;
        }
        if (p.np > 1 && p.p1[e] >= 0) {
            const double2 q = cmul(p.A[p.p1[e]], p.B[p.q1[e]]);
            v.x += p.sg1 * q.x;
            v.y += p.sg1 * q.y;
        }
        p.out[e] = v;
    }
}

void mapbil_params(double2* out, const double2* src0, const double2* A, const double2* B,
                   const int32_t* m0, const int32_t* p0, const int32_t* q0, const int32_t* p1,
                   const int32_t* q1, int64_t n, int np, double c0, double sg0, double sg1,
                   void* q) {
    *(MapBilP*)q = MapBilP{out, src0, A, B, m0, p0, q0, p1, q1, n, np, c0, sg0, sg1};
}

void k_map_bil(double2* out, const double2* src0, const double2* A, const double2* B,
               const int32_t* m0, const int32_t* p0, const int32_t* q0, const int32_t* p1,
               const int32_t* q1, int64_t n, int np, double c0, double sg0, double sg1,
               hipStream_t st) {
    ew_launch(RK_MAPBIL, MapBilP{out, src0, A, B, m0, p0, q0, p1, q1, n, np, c0, sg0, sg1}, n,
              1, st);
}

struct BDenomP { double2* S; const double* D; const double* F; int64_t n; int divide; int mask; };
__device__ __forceinline__ void body(const BDenomP& p) {
    GRID_LOOP(e, p.n) {
        double2 s = p.S[e];
        if (p.divide) {
            const double d = p.D[e];
            s.x /= d;
            s.y /= d;
        }
        if (p.mask && p.F[e] != 0.0) s = make_double2(0.0, 0.0);
        p.S[e] = s;
    }
}

void k_bdenom(double2* S, const double* D, const double* F, int64_t n, int divide, int mask,
              hipStream_t st) {
    ew_launch(RK_BDENOM, BDenomP{S, D, F, n, divide, mask}, n, 1, st);
}

struct GRedP { const GRedDesc* d; int count; };
__device__ __forceinline__ void body(const GRedP& p) {
    for (int q = blockIdx.y; q < p.count; q += gridDim.y) {
        const GRedDesc d = p.d[q];
        const int64_t mn = (int64_t)d.m * d.n;
        const int ways = split_ways(d.nsplit);
        GRID_LOOP(t, mn * ways) {
            const int64_t e = t / ways;
            const int qq = (int)(t - e * ways);
            const double2 sum = split_sum(d.W + e, d.nsplit, mn, qq, ways);
            if (qq != 0) continue;
            const int64_t i = e % d.m, j = e / d.m;
            double2* cp = d.C + i + j * d.ldc;
            double2 o = make_double2(d.alpha * sum.x, d.alpha * sum.y);
            if (d.beta != 0.0) {
                const double2 c = *cp;
                o.x += d.beta * c.x;
                o.y += d.beta * c.y;
            }
            *cp = o;
        }
    }
}

void k_gred(const GRedDesc* d, int count, int64_t maxn, hipStream_t st) {
    ew_launch(RK_GRED, GRedP{d, count}, maxn, std::max(1, std::min(count, 64)), st);
}

void gettau_params(const double2* t1, const double2* t2, double2* tau, double s2, PairTab vv,
                   PairTab oo, int nv, void* gp) {
    *(GettauP*)gp = GettauP{t1, t2, tau, s2, vv.tab, vv.npair, oo.tab, oo.npair, nv};
}

void copy_params(int64_t n, const double2* x, double2* y, void* cp) {
    *(CopyP*)cp = CopyP{n, x, y};
}

size_t batch_param_size(int kind) {
    switch (kind) {
        case RK_ZDOT_PART: return sizeof(ZdotPartP);
        case RK_FINISH: return sizeof(FinishP);
        case RK_SUMNORM_PART: return sizeof(SumNormPartP);
        case RK_LINCOMB: return sizeof(LincombP);
        case RK_DIIS_PREP: return sizeof(DiisPrepP);
        case RK_GETTAU: return sizeof(GettauP);
        case RK_COPY: return sizeof(CopyP);
        case RK_MAPBIL: return sizeof(MapBilP);
        default: return 0;
    }
}

// One launch of a finish-phase kind over `count` handles (parameters at dp,
// device): grid-stride kinds get a capped grid for the largest n, the
// reductions exactly gx partial blocks (each handle uses its own number).
int k_many(int kind, const void* dp, int count, int64_t gx, hipStream_t st) {
    if (count <= 0 || gx <= 0) return 0;
    const bool stride = kind == RK_LINCOMB || kind == RK_DIIS_PREP || kind == RK_GETTAU ||
                        kind == RK_COPY || kind == RK_MAPBIL;
    const int64_t x = stride ? std::min<int64_t>(nblocks(gx),
                                                 std::max<int64_t>(1, 65536 / count))
                             : gx;
    const dim3 grid((unsigned)x, 1, count);
#define RCC_M(K, P)                                                                        \
    case K:                                                                                \
        hipLaunchKernelGGL(ew_many<P>, grid, dim3(TPB), 0, st, (const P*)dp);              \
        break;
    switch (kind) {
        RCC_M(RK_ZDOT_PART, ZdotPartP)
        RCC_M(RK_FINISH, FinishP)
        RCC_M(RK_SUMNORM_PART, SumNormPartP)
        RCC_M(RK_LINCOMB, LincombP)
        RCC_M(RK_DIIS_PREP, DiisPrepP)
        RCC_M(RK_GETTAU, GettauP)
        RCC_M(RK_COPY, CopyP)
        RCC_M(RK_MAPBIL, MapBilP)
        default: return -1;
    }
#undef RCC_M
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ---- symmetry blocks <-> C1 layout ----
// One thread per block element (r, c), r fastest: coalesced on the packed
// block side; the C1 side is a scatter / gather through the pair maps.
__global__ void scatter_block_kernel(const double2* __restrict__ src, int64_t rows, int64_t n,
                                     const int* __restrict__ rowmap,
                                     const int* __restrict__ colmap, double2* __restrict__ dst,
                                     int64_t ld) {
    GRID_LOOP(e, n) {
        const int64_t r = e % rows, c = e / rows;
        dst[(int64_t)rowmap[r] + ld * colmap[c]] = src[e];
    }
}

__global__ void gather_block_kernel(const double2* __restrict__ src, int64_t ld,
                                    const int* __restrict__ rowmap,
                                    const int* __restrict__ colmap, int64_t rows, int64_t n,
                                    double2* __restrict__ dst) {
    GRID_LOOP(e, n) {
        const int64_t r = e % rows, c = e / rows;
        dst[e] = src[(int64_t)rowmap[r] + ld * colmap[c]];
    }
}

void k_scatter_block(const double2* src, int64_t rows, int64_t ncols, const int* rowmap,
                     const int* colmap, double2* dst, int64_t ld, hipStream_t st) {
    const int64_t n = rows * ncols;
    if (n <= 0) return;
    hipLaunchKernelGGL(scatter_block_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, src, rows, n,
                       rowmap, colmap, dst, ld);
}

void k_gather_block(const double2* src, int64_t ld, const int* rowmap, const int* colmap,
                    int64_t rows, int64_t ncols, double2* dst, hipStream_t st) {
    const int64_t n = rows * ncols;
    if (n <= 0) return;
    hipLaunchKernelGGL(gather_block_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, src, ld, rowmap,
                       colmap, rows, n, dst);
}

// All blocks of a layout in one launch: blockIdx.y = block.
struct BlockSetP { const double2* src; BlockSet bs; double2* dst; int64_t ld; int gather; };
__device__ __forceinline__ void body(const BlockSetP& p) {
    if ((int)blockIdx.y >= p.bs.nb) return;
    const BlockSet::Blk& b = p.bs.b[blockIdx.y];
    const int64_t n = b.rows * b.cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e % b.rows, c = e / b.rows;
        const int64_t c1 = (int64_t)b.rmap[r] + p.ld * b.cmap[c];
        if (p.gather) p.dst[b.off + e] = p.src[c1];
        else p.dst[c1] = p.src[b.off + e];
    }
}

void k_blockset(const double2* src, const BlockSet& bs, double2* dst, int64_t ld, bool gather,
                hipStream_t st) {
    int64_t mx = 0;
    for (int i = 0; i < bs.nb; ++i) mx = std::max<int64_t>(mx, bs.b[i].rows * bs.b[i].cols);
    ew_launch(RK_BLOCKSET, BlockSetP{src, bs, dst, ld, gather ? 1 : 0}, mx, bs.nb, st);
}

struct FillP { double2* p; int64_t n; };
__device__ __forceinline__ void body(const FillP& p) {
    GRID_LOOP(e, p.n) { p.p[e] = make_double2(0.0, 0.0); }
}

void k_fill_zero(double2* p, int64_t n, hipStream_t st) {
    ew_launch(RK_FILL, FillP{p, n}, n, 1, st);
}

// ---------------------------------------------------------------- batched replay
// The split-K reduction of zgemm as an element-wise body (sums the partials
// in split order, then alpha / beta), for replayed split products.
struct SplitReduceP { int m; int n; int nsplit; int batch; const double2* W; double2* C;
                      int64_t ldc; int64_t sC; double2 alpha; double2 beta; };
__device__ __forceinline__ void body(const SplitReduceP& p) {
    const int64_t mn = (int64_t)p.m * p.n;
    const int ways = split_ways(p.nsplit);
    GRID_LOOP(t, mn * p.batch * ways) {
        const int64_t oi = t / ways;
        const int q = (int)(t - oi * ways);
        const int64_t b = oi / mn, e = oi - b * mn;
        const int64_t i = e % p.m, j = e / p.m;
        const double2 sum = split_sum(p.W + b * p.nsplit * mn + e, p.nsplit, mn, q, ways);
        if (q != 0) continue;
        const double sr = sum.x, si = sum.y;
        double2* cp = p.C + b * p.sC + i + j * p.ldc;
        double2 o;
        o.x = p.alpha.x * sr - p.alpha.y * si;
        o.y = p.alpha.x * si + p.alpha.y * sr;
        if (p.beta.x != 0.0 || p.beta.y != 0.0) {
            const double2 c = *cp;
            o.x += p.beta.x * c.x - p.beta.y * c.y;
            o.y += p.beta.x * c.y + p.beta.y * c.x;
        }
        *cp = o;
    }
}

static size_t param_size(int kind) {
    switch (kind) {
        case RK_GETTAU: return sizeof(GettauP);
        case RK_EXPAND_E2: return sizeof(ExpandE2P);
        case RK_EXPAND_F1: return sizeof(ExpandF1P);
        case RK_EXPAND_RING: return sizeof(ExpandRingP);
        case RK_PACK_PIJ: return sizeof(PackPijP);
        case RK_PACK_PAB: return sizeof(PackPabP);
        case RK_PACK_PABIJ: return sizeof(PackPabijP);
        case RK_PERM_AKCI: return sizeof(PermAkciP);
        case RK_TRACE_T1: return sizeof(TraceT1P);
        case RK_TRACE_GV: return sizeof(TraceGvP);
        case RK_DENOM1: return sizeof(Denom1P);
        case RK_DENOM2: return sizeof(Denom2P);
        case RK_AXPBY: return sizeof(AxpbyP);
        case RK_CONJ: return sizeof(ConjP);
        case RK_COPY: return sizeof(CopyP);
        case RK_ZERO_DIAG: return sizeof(ZeroDiagP);
        case RK_BLOCKSET: return sizeof(BlockSetP);
        case RK_FILL: return sizeof(FillP);
        case RK_SPLIT_REDUCE: return sizeof(SplitReduceP);
        case RK_MAPSUM: return sizeof(MapSumP);
        case RK_MAPBIL: return sizeof(MapBilP);
        case RK_BDENOM: return sizeof(BDenomP);
        case RK_GRED: return sizeof(GRedP);
        default: return 0;
    }
}

// One element-wise kind for `count` handles, parameters at dp (device).  The
// handles' sizes differ, so the grid is not sized for the largest one times
// the count (mostly empty workgroups): at most 64K workgroups in all, each
// striding over its handle's elements.
static int ew_many_launch(int kind, const void* dp, int count, int64_t maxn, int maxgy,
                          hipStream_t st) {
    if (count <= 0 || maxn <= 0 || maxgy <= 0) return 0;
    const int64_t cap = std::max<int64_t>(1, 65536 / ((int64_t)count * maxgy));
    const dim3 grid((unsigned)std::min<int64_t>(nblocks(maxn), cap), maxgy, count);
#define RCC_EW(K, P)                                                                       \
    case K:                                                                                \
        hipLaunchKernelGGL(ew_many<P>, grid, dim3(TPB), 0, st, (const P*)dp);              \
        break;
    switch (kind) {
        RCC_EW(RK_GETTAU, GettauP)
        RCC_EW(RK_EXPAND_E2, ExpandE2P)
        RCC_EW(RK_EXPAND_F1, ExpandF1P)
        RCC_EW(RK_EXPAND_RING, ExpandRingP)
        RCC_EW(RK_PACK_PIJ, PackPijP)
        RCC_EW(RK_PACK_PAB, PackPabP)
        RCC_EW(RK_PACK_PABIJ, PackPabijP)
        RCC_EW(RK_PERM_AKCI, PermAkciP)
        RCC_EW(RK_TRACE_T1, TraceT1P)
        RCC_EW(RK_TRACE_GV, TraceGvP)
        RCC_EW(RK_DENOM1, Denom1P)
        RCC_EW(RK_DENOM2, Denom2P)
        RCC_EW(RK_AXPBY, AxpbyP)
        RCC_EW(RK_CONJ, ConjP)
        RCC_EW(RK_COPY, CopyP)
        RCC_EW(RK_ZERO_DIAG, ZeroDiagP)
        RCC_EW(RK_BLOCKSET, BlockSetP)
        RCC_EW(RK_FILL, FillP)
        RCC_EW(RK_SPLIT_REDUCE, SplitReduceP)
        RCC_EW(RK_MAPSUM, MapSumP)
        RCC_EW(RK_MAPBIL, MapBilP)
        RCC_EW(RK_BDENOM, BDenomP)
        RCC_EW(RK_GRED, GRedP)
        default: return -1;
    }
#undef RCC_EW
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

// A compiled batch: every recorded position becomes one launch over all
// handles (element-wise: one ew_many; product: one grouped GEMM over every
// handle's (batch index, k split) pieces, plus one split reduction).
struct BatchOp {
    int kind = 0;
    const void* dp = nullptr;     // element-wise parameters / grouped descriptors
    int count = 0;
    int64_t maxn = 0;
    int maxgy = 1;
    int opa = 0, opb = 0, tag = 0, gauss = 0;
    int64_t blocks = 0;
    const void* dmap = nullptr;   // grouped: block -> group map
    // grouped: the same groups numbered in 32 x 32 tiles, and which tile the
    // autotuner picked (batch_tune)
    const void* dp32 = nullptr;
    const void* dmap32 = nullptr;
    int64_t blocks32 = 0;
    // the same groups in 128 x 32 tiles
    const void* dpt = nullptr;
    int64_t blockst = 0;
    int small = 0;
    // the streaming kernel (fixed, not tuned): its descriptors and columns
    bool fixed = false;
    const void* dpsk = nullptr;
    int64_t blockssk = 0;
    int nj = 0;
    const void* dred = nullptr;   // split reductions (redcount entries) or null
    int redcount = 0;
    int64_t redn = 0;
};

struct BatchProgram {
    std::vector<BatchOp> ops;
    void* arena = nullptr;
};

static bool as_stored_op(int op) { return op == OP_N || op == OP_R; }

int batch_compile(const std::vector<std::vector<Rec>*>& recs, BatchProgram** out,
                  std::string& err) {
    *out = nullptr;
    const int nh = (int)recs.size();
    if (nh == 0) return 0;
    const size_t npos = recs[0]->size();
    for (int h = 1; h < nh; ++h)
        if (recs[h]->size() != npos) {
            err = "batch: handles record term sequences of different lengths";
            return -1;
        }
    auto* prog = new BatchProgram();
    std::vector<unsigned char> host;
    auto put = [&](const void* p, size_t n) -> size_t {
        size_t off = (host.size() + 255) / 256 * 256;
        host.resize(off + n);
        memcpy(host.data() + off, p, n);
        return off;
    };
    struct Pending { size_t off, red, map, off32, map32, offt, offsk; bool has_red; };
    std::vector<Pending> pend;
    for (size_t j = 0; j < npos; ++j) {
        const Rec& r0 = (*recs[0])[j];
        BatchOp op;
        op.kind = r0.kind;
        op.count = nh;
        Pending pd{0, 0, 0, 0, 0, 0, 0, false};
        for (int h = 1; h < nh; ++h) {
            const Rec& r = (*recs[h])[j];
            bool same = r.kind == r0.kind;
            if (same && r0.kind == RK_GEMM)
                same = r.g.opA == r0.g.opA && r.g.opB == r0.g.opB && r.g.tag == r0.g.tag &&
                       r.g.gauss == r0.g.gauss;
            if (same && r0.kind == RK_GROUPED)
                same = r.opa == r0.opa && r.opb == r0.opb && r.tag == r0.tag &&
                       r.gauss == r0.gauss;
            if (!same) {
                err = "batch: handles record different term sequences (position " +
                      std::to_string(j) + ")";
                delete prog;
                return -1;
            }
        }
        if (r0.kind == RK_GEMM || r0.kind == RK_GROUPED) {
          // handles whose plans put this product on different kernels (the
          // streaming one or the MFMA tiles) get one launch per kernel
          auto vof = [&](int h) {
              const Rec& r = (*recs[h])[j];
              return r.kind == RK_GROUPED ? r.variant : -1;
          };
          std::vector<int> vset;
          for (int h = 0; h < nh; ++h)
              if (std::find(vset.begin(), vset.end(), vof(h)) == vset.end()) vset.push_back(vof(h));
          const BatchOp op_init = op;
          const Pending pd_init = pd;
          for (size_t vi = 0; vi < vset.size(); ++vi) {
            const int want = vset[vi];
            if (vi > 0) {
                op = op_init;
                pd = pd_init;
            }
            std::vector<ZGroup> all;
            std::vector<SplitReduceP> red(nh);
            bool any_red = false;
            // K splits were chosen to fill the chip with one handle; over many
            // handles fewer (longer) splits do: scale them down until the
            // launch has about 16K workgroups
            std::vector<int> ns(nh, 1), kc(nh, 0);
            int64_t total = 0;
            for (int h = 0; h < nh; ++h) {
                if (vof(h) != want) continue;
                const Rec& r = (*recs[h])[j];
                ns[h] = r.nsplit;
                kc[h] = r.kchunk;
                if (r.kind != RK_GEMM) continue;
                const int64_t tl = (int64_t)((r.g.m + 63) / 64) * ((r.g.n + 63) / 64);
                total += tl * std::max(r.g.batch, 0) * r.nsplit;
            }
            const int64_t f = (total + 16383) / 16384;
            for (int h = 0; h < nh && f > 1; ++h) {
                if (vof(h) != want) continue;
                const Rec& r = (*recs[h])[j];
                if (r.kind != RK_GEMM || r.nsplit <= 1 || r.g.k <= 0) continue;
                const int want = (int)std::max<int64_t>(1, (r.nsplit + f - 1) / f);
                int c = (r.g.k + want - 1) / want;
                c = (c + 15) / 16 * 16;
                ns[h] = (r.g.k + c - 1) / c;
                kc[h] = ns[h] > 1 ? c : r.g.k;
            }
            for (int h = 0; h < nh; ++h) {
                if (vof(h) != want) continue;
                Rec r = (*recs[h])[j];
                r.nsplit = ns[h];
                r.kchunk = kc[h];
                red[h] = SplitReduceP{0, 0, 1, 0, nullptr, nullptr, 1, 0, {0, 0}, {0, 0}};
                if (r.kind == RK_GROUPED) {
                    all.insert(all.end(), r.groups.begin(), r.groups.end());
                    continue;
                }
                const GemmArgs& g = r.g;
                if (g.m <= 0 || g.n <= 0 || g.batch <= 0) continue;
                if (g.k <= 0) {
                    err = "batch: a product with k = 0 has no grouped form";
                    delete prog;
                    return -1;
                }
                for (int b = 0; b < g.batch; ++b)
                    for (int s = 0; s < r.nsplit; ++s) {
                        const int kb = s * r.kchunk;
                        const int kl = std::min(r.kchunk, g.k - kb);
                        if (kl <= 0) continue;
                        ZGroup z{};
                        z.m = g.m; z.n = g.n; z.k = kl;
                        z.A = g.A + b * g.sA + (as_stored_op(g.opA) ? (int64_t)kb * g.lda : kb);
                        z.lda = g.lda;
                        z.B = g.B + b * g.sB + (as_stored_op(g.opB) ? kb : (int64_t)kb * g.ldb);
                        z.ldb = g.ldb;
                        if (r.nsplit > 1) {
                            z.C = r.ws + ((int64_t)b * r.nsplit + s) * g.m * g.n;
                            z.ldc = g.m;
                            z.alpha = make_double2(1.0, 0.0);
                            z.beta = make_double2(0.0, 0.0);
                        } else {
                            z.C = g.C + b * g.sC;
                            z.ldc = g.ldc;
                            z.alpha = g.alpha;
                            z.beta = g.beta;
                        }
                        all.push_back(z);
                    }
                if (r.nsplit > 1) {
                    red[h] = SplitReduceP{g.m, g.n, r.nsplit, g.batch, r.ws, g.C, g.ldc, g.sC,
                                          g.alpha, g.beta};
                    any_red = true;
                    op.redn = std::max<int64_t>(op.redn, (int64_t)g.m * g.n * g.batch);
                }
            }
            op.kind = RK_GROUPED;
            op.opa = r0.kind == RK_GEMM ? r0.g.opA : r0.opa;
            op.opb = r0.kind == RK_GEMM ? r0.g.opB : r0.opb;
            op.tag = r0.kind == RK_GEMM ? r0.g.tag : r0.tag;
            op.gauss = r0.kind == RK_GEMM ? r0.g.gauss : r0.gauss;
            op.count = (int)all.size();
            op.blocks = zgroup_offsets(all.data(), op.count);
            if (!all.empty()) {
                pd.off = put(all.data(), all.size() * sizeof(ZGroup));
                std::vector<int> gm((size_t)op.blocks, 0);
                for (int q = 0; q < op.count; ++q) {
                    const int64_t end = q + 1 < op.count ? all[q + 1].tile_off : op.blocks;
                    for (int64_t b = all[q].tile_off; b < end; ++b) gm[(size_t)b] = q;
                }
                pd.map = put(gm.data(), gm.size() * sizeof(int));
                std::vector<ZGroup> s32 = all;
                op.blocks32 = zgroup_offsets_small(s32.data(), op.count);
                pd.off32 = put(s32.data(), s32.size() * sizeof(ZGroup));
                std::vector<int> gm32((size_t)op.blocks32, 0);
                for (int q = 0; q < op.count; ++q) {
                    const int64_t end = q + 1 < op.count ? s32[q + 1].tile_off : op.blocks32;
                    for (int64_t b = s32[q].tile_off; b < end; ++b) gm32[(size_t)b] = q;
                }
                pd.map32 = put(gm32.data(), gm32.size() * sizeof(int));
                std::vector<ZGroup> st = all;
                op.blockst = zgroup_offsets_tall(st.data(), op.count);
                pd.offt = put(st.data(), st.size() * sizeof(ZGroup));
                if (want == 3) {
                    std::vector<ZGroup> sk = all;
                    op.blockssk = zgroup_offsets_skinny(sk.data(), op.count);
                    pd.offsk = put(sk.data(), sk.size() * sizeof(ZGroup));
                    op.fixed = true;
                    op.small = 3;
                    for (int h = 0; h < nh; ++h)
                        if (vof(h) == 3) op.nj = std::max(op.nj, (*recs[h])[j].nj);
                }
            }
            if (any_red) {
                pd.red = put(red.data(), red.size() * sizeof(SplitReduceP));
                pd.has_red = true;
                op.redcount = nh;
            }
            if (vi + 1 < vset.size()) {
                prog->ops.push_back(op);
                pend.push_back(pd);
            }
          }
        } else {
            const size_t ps = param_size(r0.kind);
            if (ps == 0) {
                err = "batch: a recorded launch has no batched form";
                delete prog;
                return -1;
            }
            std::vector<unsigned char> blob((size_t)nh * ps);
            for (int h = 0; h < nh; ++h) {
                const Rec& r = (*recs[h])[j];
                memcpy(blob.data() + (size_t)h * ps, r.blob.data(), ps);
                op.maxn = std::max<int64_t>(op.maxn, r.n);
                op.maxgy = std::max(op.maxgy, r.gy);
            }
            pd.off = put(blob.data(), blob.size());
        }
        prog->ops.push_back(op);
        pend.push_back(pd);
    }
    if (!host.empty()) {
        if (hipMalloc(&prog->arena, host.size()) != hipSuccess ||
            hipMemcpy(prog->arena, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) {
            err = "batch: parameter upload failed";
            if (prog->arena) (void)hipFree(prog->arena);
            delete prog;
            return -1;
        }
    }
    for (size_t j = 0; j < prog->ops.size(); ++j) {
        prog->ops[j].dp = (const unsigned char*)prog->arena + pend[j].off;
        if (prog->ops[j].kind == RK_GROUPED && prog->ops[j].count > 0) {
            prog->ops[j].dmap = (const unsigned char*)prog->arena + pend[j].map;
            prog->ops[j].dp32 = (const unsigned char*)prog->arena + pend[j].off32;
            prog->ops[j].dmap32 = (const unsigned char*)prog->arena + pend[j].map32;
            prog->ops[j].dpt = (const unsigned char*)prog->arena + pend[j].offt;
            if (prog->ops[j].fixed)
                prog->ops[j].dpsk = (const unsigned char*)prog->arena + pend[j].offsk;
        }
        if (pend[j].has_red) prog->ops[j].dred = (const unsigned char*)prog->arena + pend[j].red;
    }
    *out = prog;
    return 0;
}

// Kernels of a grouped product: 0 the 96 x 64 tile, 1 the 32 x 32 tile (one
// workgroup per tile), 2 the 32 x 32 tile on a persistent grid of 1024
// workgroups (four per CU) striding over the tiles, 3 the streaming kernel
// (fixed by the plan, not tuned), 4 the 128 x 32 tile.  The MFMA tiles give
// the same bits.
constexpr int TUNED_VARIANTS[] = {0, 1, 2, 4};
constexpr int NTUNED = 4;

static int grouped_launch(const BatchOp& op, int small, hipStream_t st) {
    if (op.count <= 0) return 0;
    if (small == 3)
        return zgemm_grouped_skinny(op.nj, (const ZGroup*)op.dpsk, op.count, op.blockssk, st);
    if (small == 4)
        return zgemm_grouped_tall(op.opa, op.opb, op.gauss, (const ZGroup*)op.dpt, op.count,
                                  op.blockst, st);
    if (small)
        return zgemm_grouped_small(op.opa, op.opb, op.gauss, (const ZGroup*)op.dp32, op.count,
                                   op.blocks32, st, (const int*)op.dmap32,
                                   small == 2 ? 1024 : 0);
    return zgemm_grouped_ops(op.opa, op.opb, op.tag, op.gauss, (const ZGroup*)op.dp, op.count,
                             op.blocks, st, (const int*)op.dmap);
}

// Tile autotuner of a compiled batch: every grouped product is timed with
// each tile variant (HIP events, best of two runs after a warm-up) and keeps
// the fastest.  Run before the first replay: the products
// then write only intermediates that the term sequence recomputes from T.
int batch_tune(BatchProgram* p, hipStream_t st) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1;
    for (BatchOp& op : p->ops) {
        if (op.kind != RK_GROUPED || op.count <= 0 || op.fixed) continue;
        float best[NTUNED] = {1e30f, 1e30f, 1e30f, 1e30f};
        for (int v = 0; v < NTUNED; ++v)
            for (int rep = 0; rep < 3; ++rep) {
                (void)hipEventRecord(e0, st);
                if (grouped_launch(op, TUNED_VARIANTS[v], st) != 0) {
                    (void)hipGetLastError();
                    best[v] = 1e30f;
                    break;
                }
                (void)hipEventRecord(e1, st);
                (void)hipEventSynchronize(e1);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (rep > 0) best[v] = std::min(best[v], ms);
            }
        int pick = 0;
        for (int v = 1; v < NTUNED; ++v)
            if (best[v] < best[pick]) pick = v;
        op.small = TUNED_VARIANTS[pick];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

void batch_info(const BatchProgram* p, int* positions, int* grouped, int* variants) {
    int g = 0;
    for (int v = 0; v < 5; ++v) variants[v] = 0;
    for (const BatchOp& op : p->ops)
        if (op.kind == RK_GROUPED && op.count > 0) {
            ++g;
            variants[std::min(std::max(op.small, 0), 4)] += 1;
        }
    *positions = (int)p->ops.size();
    *grouped = g;
}

int batch_run(const BatchProgram* p, hipStream_t st) {
    for (const BatchOp& op : p->ops) {
        if (op.kind == RK_GROUPED) {
            if (grouped_launch(op, op.small, st) != 0) return -1;
            if (op.dred &&
                ew_many_launch(RK_SPLIT_REDUCE, op.dred, op.redcount, op.redn, 1, st) != 0)
                return -1;
        } else {
            if (ew_many_launch(op.kind, op.dp, op.count, op.maxn, op.maxgy, st) != 0) return -1;
        }
    }
    return 0;
}

void batch_free(BatchProgram* p) {
    if (!p) return;
    if (p->arena) (void)hipFree(p->arena);
    delete p;
}

}  // namespace rcc
