// Complex double-precision GEMM on CDNA4 FP64 matrix cores (gfx950).
//
//   C(m x n) = alpha * op(A)(m x k) * op(B)(k x n) + beta * C      (column-major)
//
// op is N (as stored), T (transposed), C (conjugate-transposed) or R
// (conjugated, not transposed; with the other operand N).  Operands are
// interleaved complex (re, im) doubles, the layout relccsd's REAL*8 pairs use
// (RCW = 2).  The tile machinery -- direct global -> LDS copies of 8-deep K
// stages into two LDS buffers, interleaved complex in LDS, the 3M or 4M
// complex product on v_mfma_f64_16x16x4_f64 -- is in rcc_zgemm_core.h and is
// shared with the triples GEMM.
//
// Work decomposition: a 256-thread workgroup (2 x 2 wave64) owns a 96 x 64
// complex C tile, each wave 48 x 32 (3 x 2 MFMA blocks of 16 x 16); 40 KB
// of LDS, so two workgroups share a CU and each one's barrier waits are
// covered by the other's MFMAs.  Tiles are numbered so that each of the 8
// XCDs (workgroups are dealt to XCDs round-robin) owns a contiguous range of
// tiles ordered N-fastest: the workgroups sharing one M-panel of A run
// together on one XCD and read it through that XCD's L2 once.  grid.z
// enumerates (batch, K-split).  K-splits write raw partial sums to a
// workspace that a second kernel reduces in a fixed order, so results are
// deterministic.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>
#include "rcc_internal.h"
#include "rcc_zgemm_core.h"

namespace rcc {

// The default tile, and a skinny one (128 x 32, waves of 64 x 16) for
// products with n <= 32: the occupied-index products (n = o) would leave
// half of a 64-wide tile idle.
using TLS = ZTile<4, 1>;
constexpr int BM = ZTileGemm::BM;
constexpr int BN = ZTileGemm::BN;
constexpr int ZBAND = 8;      // tile rows per band of the tile order

// the tile of a product: skinny for untagged products with n <= 32 (the
// tagged ladder / ring launches and the grouped launches use the default)
constexpr bool skinny_n(int n) { return n <= TLS::BN; }
int64_t tile_count(const GemmArgs& g) {
    const bool sk = g.tag == 0 && skinny_n(g.n);
    const int bm = sk ? TLS::BM : BM, bn = sk ? TLS::BN : BN;
    return (int64_t)((g.m + bm - 1) / bm) * ((g.n + bn - 1) / bn) * g.batch;
}

template <int OP>
__device__ __forceinline__ double2 cj(double2 v) {
    if (OP == OP_C || OP == OP_R) v.y = -v.y;
    return v;
}

// N and R read the operand as stored (R conjugates), T and C transposed
template <int OP>
constexpr bool as_stored() { return OP == OP_N || OP == OP_R; }
template <int OP>
constexpr bool conj_op() { return OP == OP_C || OP == OP_R; }

// One workgroup: one TL tile of C for batch index / K-split given by blockIdx.z.
// split == 1: C = alpha * acc + beta * C.  split > 1: raw acc goes to
// W[(z) * m * n + i + j * m] and zgemm_reduce finishes the epilogue.

// GROUPED: the problem (m, n, k, operands, alpha, beta) is read from the
// group descriptor that owns this block (binary search on tile offsets); no
// batch, no K-split.
template <int OPA, int OPB, int TAG, int GAUSS, int GROUPED = 0, class TL = ZTileGemm>
__global__ void __launch_bounds__(256, 2)     // two workgroups (two waves per SIMD) per CU
zgemm_mfma_kernel(int m, int n, int k,
                  const double2* __restrict__ A, int64_t lda, int64_t sA,
                  const double2* __restrict__ B, int64_t ldb, int64_t sB,
                  double2* __restrict__ C, int64_t ldc, int64_t sC,
                  double2 alpha, double2 beta, int nsplit, int kchunk,
                  double2* __restrict__ W, const ZGroup* __restrict__ G = nullptr,
                  int ngroups = 0, const int* __restrict__ gmap = nullptr, int band_rows = ZBAND) {
    // two stage buffers, distinct arrays (see rcc_zgemm_core.h)
    __shared__ double2 lds0[TL::SA + TL::SB];
    __shared__ double2 lds1[TL::SA + TL::SB];
    constexpr int BM = TL::BM, BN = TL::BN;
    constexpr bool AR = as_stored<OPA>();     // A: m fastest in memory
    constexpr bool BR = !as_stored<OPB>();    // B (k x n as used): n fastest in memory

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    // XCD-aware tile order (see header): block b runs on XCD b % 8
    int b = blockIdx.x;
    if (GROUPED) {
        // the owning group: from the block -> group map when given (launches
        // over thousands of small groups), else by binary search on tile_off
        int lo = 0, hi = ngroups - 1;
        if (gmap) lo = gmap[b];
        else
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (G[mid].tile_off <= (int64_t)b) lo = mid;
                else hi = mid - 1;
            }
        const ZGroup& gg = G[lo];
        m = gg.m; n = gg.n; k = gg.k;
        A = gg.A; lda = gg.lda; sA = 0;
        B = gg.B; ldb = gg.ldb; sB = 0;
        C = gg.C; ldc = gg.ldc; sC = 0;
        alpha = gg.alpha; beta = gg.beta;
        nsplit = 1;
        kchunk = k;
        b -= (int)gg.tile_off;
        // padding blocks of the group (whole workgroup, before any barrier)
        if (b >= ((m + BM - 1) / BM) * ((n + BN - 1) / BN)) return;
    }
    const int mt = (m + BM - 1) / BM, nt = (n + BN - 1) / BN;
    const int T = mt * nt;
    const int xcd = b % NXCD, idx = b / NXCD;
    const int q = T / NXCD, rr = T % NXCD;
    const int start = xcd < rr ? xcd * (q + 1) : rr * (q + 1) + (xcd - rr) * q;
    const int z = blockIdx.z;
    const int batch = z / nsplit;
    const int split = z - batch * nsplit;
    A += (int64_t)batch * sA;
    B += (int64_t)batch * sB;
    const int kbeg = split * kchunk;
    const int kend = min(k, kbeg + kchunk);
    const int64_t ka = AR ? lda : 1, kb = BR ? ldb : 1;
    const double2* A0 = A + (int64_t)kbeg * ka;
    const double2* B0 = B + (int64_t)kbeg * kb;
    const int nst = kend > kbeg ? (kend - kbeg) / ZBK : 0;

    {
        // bands of band_rows tile rows, column by column inside a band: the
        // workgroups an XCD runs at once cover a near-square block of tiles
        // and share its A and B panels through the XCD's L2 (band_rows = 1:
        // N-fastest rows)
        int tm, tn;
        {
            const int t = start + idx;
            const int band = band_rows * nt;
            const int first = (t / band) * band_rows;
            const int rows = min(mt - first, band_rows);
            const int r = t - (t / band) * band;
            tm = first + r % rows;
            tn = r / rows;
        }
        const int m0 = tm * BM;
        const int n0 = tn * BN;
        // per-lane element offsets of this thread's copies, relative to the
        // stage base (constant over k); rows past the matrix edge are clamped to
        // the last one: they only feed outputs that are never stored
        int64_t offa[TL::LA], offb[TL::LB];
#pragma unroll
        for (int j = 0; j < TL::LA; ++j) {
            int r, kq;
            zslot<AR, BM>((j * 4 + wave) * 64 + lane, r, kq);
            const int64_t g = min(m0 + r, m - 1);
            offa[j] = AR ? g + (int64_t)kq * lda : (int64_t)kq + g * lda;
        }
#pragma unroll
        for (int j = 0; j < TL::LB; ++j) {
            int r, kq;
            zslot<BR, BN>((j * 4 + wave) * 64 + lane, r, kq);
            const int64_t g = min(n0 + r, n - 1);
            offb[j] = BR ? g + (int64_t)kq * ldb : (int64_t)kq + g * ldb;
        }

        ZAcc<TL> acc;
        acc.zero();
        zloop<TL, GAUSS>(
            acc, nst, lds0, lds1,
            [&](int s, double2* buf) {
                const double2* ga = A0 + (int64_t)s * ZBK * ka;
                const double2* gb = B0 + (int64_t)s * ZBK * kb;
#pragma unroll
                for (int j = 0; j < TL::LA; ++j) zcopy(ga + offa[j], buf + (j * 4 + wave) * 64);
#pragma unroll
                for (int j = 0; j < TL::LB; ++j)
                    zcopy(gb + offb[j], buf + TL::SA + (j * 4 + wave) * 64);
            },
            [&](int, double2* buf, ZOps<TL>& o) {
                zread<TL, AR, BR>(o, buf, buf + TL::SA, wm, wn, lane, conj_op<OPA>(), conj_op<OPB>());
            });
        // the last, partial stage: register copies, zero past kend
        const int kt = kbeg + nst * ZBK;
        if (kt < kend) {
            __syncthreads();
            double2* buf = (nst & 1) ? lds1 : lds0;
            for (int p = tid; p < TL::SA; p += ZNT) {
                int r, kq;
                zslot<AR, BM>(p, r, kq);
                const int64_t g = min(m0 + r, m - 1);
                double2 v = make_double2(0.0, 0.0);
                if (kt + kq < kend) v = AR ? A[g + (int64_t)(kt + kq) * lda] : A[(kt + kq) + g * lda];
                buf[p] = v;
            }
            for (int p = tid; p < TL::SB; p += ZNT) {
                int r, kq;
                zslot<BR, BN>(p, r, kq);
                const int64_t g = min(n0 + r, n - 1);
                double2 v = make_double2(0.0, 0.0);
                if (kt + kq < kend) v = BR ? B[g + (int64_t)(kt + kq) * ldb] : B[(kt + kq) + g * ldb];
                buf[TL::SA + p] = v;
            }
            __syncthreads();
            ZOps<TL> o;
            zread<TL, AR, BR>(o, buf, buf + TL::SA, wm, wn, lane, conj_op<OPA>(), conj_op<OPB>());
            zmma<TL, GAUSS>(acc, o);
        }

#pragma unroll
        for (int i = 0; i < TL::TI; ++i)
#pragma unroll
            for (int j = 0; j < TL::TJ; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int gm = m0 + zrow<TL>(wm, i, lane, r);
                    const int gn = n0 + zcol<TL>(wn, j, lane);
                    if (gm < m && gn < n) {
                        const double2 x = acc.template value<GAUSS>(i, j, r);
                        if (nsplit > 1) {
                            W[(int64_t)z * m * n + gm + (int64_t)gn * m] = x;
                        } else {
                            double2* cp = C + (int64_t)batch * sC + gm + (int64_t)gn * ldc;
                            double2 o;
                            o.x = alpha.x * x.x - alpha.y * x.y;
                            o.y = alpha.x * x.y + alpha.y * x.x;
                            if (beta.x != 0.0 || beta.y != 0.0) {
                                const double2 c = *cp;
                                o.x += beta.x * c.x - beta.y * c.y;
                                o.y += beta.x * c.y + beta.y * c.x;
                            }
                            *cp = o;
                        }
                    }
                }
    }
}

// Small-tile grouped GEMM: a 32 x 32 C tile per workgroup (four waves, one
// 16 x 16 MFMA block each), one LDS stage filled from registers that were
// loaded a k-step ahead.  For groups much smaller than the
// 96 x 64 tile (the per-irrep blocks of small systems: NOOT, o or v by
// irrep), where the large tile would mostly compute padding.  Tiles of a
// group are numbered from tile_off as 32 x 32 tiles (zgroup_offsets_small).
constexpr int SB = 32;
constexpr int SBK = 16;     // k per step of the small-tile kernel

template <int OPA, int OPB, int GAUSS>
__device__ __forceinline__ void small_tile(int b, const ZGroup* __restrict__ G, int ngroups,
                                           const int* __restrict__ gmap,
                                           double (&As_re)[SBK][SB + 1],
                                           double (&As_im)[SBK][SB + 1],
                                           double (&Bs_re)[SBK][SB + 1],
                                           double (&Bs_im)[SBK][SB + 1]) {
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int lo = 0;
    if (gmap) {
        lo = gmap[b];
    } else {
        int hi = ngroups - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (G[mid].tile_off <= (int64_t)b) lo = mid;
            else hi = mid - 1;
        }
    }
    const ZGroup& gg = G[lo];
    const int m = gg.m, n = gg.n, k = gg.k;
    b -= (int)gg.tile_off;
    const int nt = (n + SB - 1) / SB;
    if (b >= ((m + SB - 1) / SB) * nt) return;      // whole workgroup, before any barrier
    const int m0 = (b / nt) * SB, n0 = (b % nt) * SB;
    const double2* __restrict__ A = gg.A;
    const double2* __restrict__ Bm = gg.B;
    v4d acc0 = (v4d){0.0, 0.0, 0.0, 0.0}, acc1 = acc0, acc2 = acc0;
    const int lr = lane & 15, lk = lane >> 4;
    // this thread's two A and two B elements of a k-step (k0): read into
    // registers one step ahead, so the next step's loads are in flight while
    // the MFMAs of the current one run
    int ar[2], ak[2], bc[2], bk[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if (as_stored<OPA>()) { ar[j] = tid & 31; ak[j] = (tid >> 5) + 8 * j; }
        else { ak[j] = tid & 15; ar[j] = (tid >> 4) + 16 * j; }
        if (as_stored<OPB>()) { bk[j] = tid & 15; bc[j] = (tid >> 4) + 16 * j; }
        else { bc[j] = tid & 31; bk[j] = (tid >> 5) + 8 * j; }
    }
    double2 ra[2], rb[2];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int gm = min(m0 + ar[j], m - 1), gka = k0 + ak[j];
            ra[j] = make_double2(0.0, 0.0);
            if (gka < k)
                ra[j] = cj<OPA>(as_stored<OPA>() ? A[gm + (int64_t)gka * gg.lda]
                                                 : A[gka + (int64_t)gm * gg.lda]);
            const int gn = min(n0 + bc[j], n - 1), gkb = k0 + bk[j];
            rb[j] = make_double2(0.0, 0.0);
            if (gkb < k)
                rb[j] = cj<OPB>(as_stored<OPB>() ? Bm[gkb + (int64_t)gn * gg.ldb]
                                                 : Bm[gn + (int64_t)gkb * gg.ldb]);
        }
    };
    if (k > 0) fetch(0);
    for (int k0 = 0; k0 < k; k0 += SBK) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            As_re[ak[j]][ar[j]] = ra[j].x;
            As_im[ak[j]][ar[j]] = ra[j].y;
            Bs_re[bk[j]][bc[j]] = rb[j].x;
            Bs_im[bk[j]][bc[j]] = rb[j].y;
        }
        __syncthreads();
        if (k0 + SBK < k) fetch(k0 + SBK);
#pragma unroll
        for (int kk = 0; kk < SBK; kk += 4) {
            const int kr = kk + lk;
            const double are = As_re[kr][wm * 16 + lr], aim = As_im[kr][wm * 16 + lr];
            const double bre = Bs_re[kr][wn * 16 + lr], bim = Bs_im[kr][wn * 16 + lr];
            if (GAUSS) {
                acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(are + aim, bre + bim, acc1, 0, 0, 0);
                acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bim, acc2, 0, 0, 0);
            } else {
                acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim, acc1, 0, 0, 0);
                acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, -bim, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre, acc1, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // D: lane l, element r -> row l/16 + 4 r, column l%16 of the wave's block
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm * 16 + lk + 4 * r;
        const int gn = n0 + wn * 16 + lr;
        if (gm < m && gn < n) {
            double xr, xi;
            if (GAUSS) {
                xr = acc0[r] - acc2[r];
                xi = acc1[r] - acc0[r] - acc2[r];
            } else {
                xr = acc0[r];
                xi = acc1[r];
            }
            double2* cp = gg.C + gm + (int64_t)gn * gg.ldc;
            double2 o;
            o.x = gg.alpha.x * xr - gg.alpha.y * xi;
            o.y = gg.alpha.x * xi + gg.alpha.y * xr;
            if (gg.beta.x != 0.0 || gg.beta.y != 0.0) {
                const double2 c = *cp;
                o.x += gg.beta.x * c.x - gg.beta.y * c.y;
                o.y += gg.beta.x * c.y + gg.beta.y * c.x;
            }
            *cp = o;
        }
    }
}

// One tile per workgroup, or (persistent) a grid of at most a few workgroups
// per CU that strides over the tiles: tile b, b + gridDim.x, ...  (gridDim.x
// a multiple of 8, so a workgroup stays on its XCD's tiles).
template <int OPA, int OPB, int GAUSS>
__global__ void __launch_bounds__(256)
zgemm_small_grouped_kernel(const ZGroup* __restrict__ G, int ngroups,
                           const int* __restrict__ gmap, int64_t blocks) {
    __shared__ double As_re[SBK][SB + 1], As_im[SBK][SB + 1];
    __shared__ double Bs_re[SBK][SB + 1], Bs_im[SBK][SB + 1];
    for (int64_t b = blockIdx.x; b < blocks; b += gridDim.x) {
        small_tile<OPA, OPB, GAUSS>((int)b, G, ngroups, gmap, As_re, As_im, Bs_re, Bs_im);
        __syncthreads();     // the next tile's LDS writes after every wave's last reads
    }
}

int64_t zgroup_offsets_small(ZGroup* g, int ngroups) {
    int64_t off = 0;
    for (int i = 0; i < ngroups; ++i) {
        g[i].tile_off = off;
        off += (int64_t)((g[i].m + SB - 1) / SB) * ((g[i].n + SB - 1) / SB);
    }
    return off;
}

int zgemm_grouped_small(int opa, int opb, int gauss, const ZGroup* dG, int ngroups,
                        int64_t blocks, hipStream_t st, const int* gmap, int persistent) {
    if (ngroups <= 0 || blocks <= 0) return 0;
    const unsigned grid =
        (unsigned)(persistent ? std::min<int64_t>(blocks, (int64_t)persistent) : blocks);
#define RCC_S(a, b)                                                                            \
    if (opa == a && opb == b) {                                                                \
        if (gauss)                                                                             \
            hipLaunchKernelGGL((zgemm_small_grouped_kernel<a, b, 1>), dim3(grid), dim3(256), 0, \
                               st, dG, ngroups, gmap, blocks);                                 \
        else                                                                                   \
            hipLaunchKernelGGL((zgemm_small_grouped_kernel<a, b, 0>), dim3(grid), dim3(256), 0, \
                               st, dG, ngroups, gmap, blocks);                                 \
        return hipGetLastError() == hipSuccess ? 0 : -2;                                        \
    }
    RCC_S(OP_N, OP_N)
    RCC_S(OP_N, OP_C)
    RCC_S(OP_N, OP_T)
    RCC_S(OP_T, OP_N)
    RCC_S(OP_C, OP_N)
#undef RCC_S
    return -1;
}

// Sums the K-split partials (split_sum: split_ways(nsplit) lanes per output)
// and applies alpha / beta.
__global__ void zgemm_reduce_kernel(int m, int n, int nsplit, int batch,
                                    const double2* __restrict__ W,
                                    double2* __restrict__ C, int64_t ldc, int64_t sC,
                                    double2 alpha, double2 beta) {
    const int64_t mn = (int64_t)m * n;
    const int ways = split_ways(nsplit);
    const int64_t total = mn * batch * ways;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t oi = t / ways;
        const int q = (int)(t - oi * ways);
        const int64_t b = oi / mn;
        const int64_t e = oi - b * mn;
        const int64_t i = e % m, j = e / m;
        const double2 sum = split_sum(W + b * nsplit * mn + e, nsplit, mn, q, ways);
        if (q != 0) continue;
        const double sr = sum.x, si = sum.y;
        double2* cp = C + b * sC + i + j * ldc;
        double2 o;
        o.x = alpha.x * sr - alpha.y * si;
        o.y = alpha.x * si + alpha.y * sr;
        if (beta.x != 0.0 || beta.y != 0.0) {
            const double2 c = *cp;
            o.x += beta.x * c.x - beta.y * c.y;
            o.y += beta.x * c.y + beta.y * c.x;
        }
        *cp = o;
    }
}

__global__ void zscale_kernel(int m, int n, int batch, double2* __restrict__ C, int64_t ldc,
                              int64_t sC, double2 beta) {
    const int64_t mn = (int64_t)m * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < mn * batch;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = t / mn, e = t - b * mn;
        double2* cp = C + b * sC + (e % m) + (e / m) * ldc;
        const double2 c = *cp;
        double2 o;
        o.x = beta.x * c.x - beta.y * c.y;
        o.y = beta.x * c.y + beta.y * c.x;
        if (beta.x == 0.0 && beta.y == 0.0) o = make_double2(0.0, 0.0);
        *cp = o;
    }
}

// workgroups of the engine resident at once on the current device (two per CU)
int64_t resident_slots() {
    static int64_t slots[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 512;
    if (!slots[dev]) {
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cu <= 0)
            cu = 256;
        slots[dev] = 2 * (int64_t)cu;
    }
    return slots[dev];
}


template <int OPA, int OPB, int TAG, class TLX>
static void launch_tl(const GemmArgs& g, int nsplit, int kchunk, double2* W, hipStream_t st) {
    const int64_t tiles = (int64_t)((g.m + TLX::BM - 1) / TLX::BM) * ((g.n + TLX::BN - 1) / TLX::BN);
    dim3 grid((unsigned)tiles, 1, g.batch * nsplit);
    // the ladder's few tile columns share its A panels best row by row
    // (measured: fewer fetched bytes); the rest in bands of 8 tile rows
    const int band = TAG == 1 ? 1 : ZBAND;
    if (g.gauss)
        hipLaunchKernelGGL((zgemm_mfma_kernel<OPA, OPB, TAG, 1, 0, TLX>), grid, dim3(256), 0, st,
                           g.m, g.n, g.k, g.A, g.lda, g.sA, g.B, g.ldb, g.sB, g.C, g.ldc, g.sC,
                           g.alpha, g.beta, nsplit, kchunk, W, nullptr, 0, nullptr, band);
    else
        hipLaunchKernelGGL((zgemm_mfma_kernel<OPA, OPB, TAG, 0, 0, TLX>), grid, dim3(256), 0, st,
                           g.m, g.n, g.k, g.A, g.lda, g.sA, g.B, g.ldb, g.sB, g.C, g.ldc, g.sC,
                           g.alpha, g.beta, nsplit, kchunk, W, nullptr, 0, nullptr, band);
}

template <int OPA, int OPB, int TAG>
static void launch_t(const GemmArgs& g, int nsplit, int kchunk, double2* W, hipStream_t st) {
    if (TAG == 0 && skinny_n(g.n)) launch_tl<OPA, OPB, TAG, TLS>(g, nsplit, kchunk, W, st);
    else launch_tl<OPA, OPB, TAG, ZTileGemm>(g, nsplit, kchunk, W, st);
}

template <int OPA, int OPB>
static void launch(const GemmArgs& g, int nsplit, int kchunk, double2* W, hipStream_t st) {
    // tagged symbols: the VVVV ladder and its adjoint (1), the (ov)^3 rings
    // and their adjoints (2)
    constexpr bool ladder = OPB == OP_N && (OPA == OP_N || OPA == OP_T);
    constexpr bool ring = OPA == OP_N || (OPA == OP_T && OPB == OP_N);
    if (ladder && g.tag == 1) launch_t<OPA, OPB, ladder ? 1 : 0>(g, nsplit, kchunk, W, st);
    else if (ring && g.tag == 2) launch_t<OPA, OPB, ring ? 2 : 0>(g, nsplit, kchunk, W, st);
    else launch_t<OPA, OPB, 0>(g, nsplit, kchunk, W, st);
}

int64_t zgroup_offsets(ZGroup* g, int ngroups) {
    int64_t off = 0;
    for (int i = 0; i < ngroups; ++i) {
        g[i].tile_off = off;
        const int64_t t = (int64_t)((g[i].m + BM - 1) / BM) * ((g[i].n + BN - 1) / BN);
        off += (t + NXCD - 1) / NXCD * NXCD;
    }
    return off;
}

template <int OPB, int TAG, int GAUSS>
static void launch_grouped(const ZGroup* dG, int ngroups, int64_t blocks, hipStream_t st,
                           const int* gmap) {
    hipLaunchKernelGGL((zgemm_mfma_kernel<OP_N, OPB, TAG, GAUSS, 1>), dim3((unsigned)blocks),
                       dim3(256), 0, st, 0, 0, 0, (const double2*)nullptr, (int64_t)0, (int64_t)0,
                       (const double2*)nullptr, (int64_t)0, (int64_t)0, (double2*)nullptr,
                       (int64_t)0, (int64_t)0, make_double2(0, 0), make_double2(0, 0), 1, 0,
                       (double2*)nullptr, dG, ngroups, gmap);
}

int zgemm_grouped(int opb, int tag, int gauss, const ZGroup* dG, int ngroups, int64_t blocks,
                  hipStream_t st, const int* gmap) {
    if (ngroups <= 0 || blocks <= 0) return 0;
    if (opb == OP_N) {
        if (tag == 1) gauss ? launch_grouped<OP_N, 1, 1>(dG, ngroups, blocks, st, gmap)
                            : launch_grouped<OP_N, 1, 0>(dG, ngroups, blocks, st, gmap);
        else if (tag == 2) gauss ? launch_grouped<OP_N, 2, 1>(dG, ngroups, blocks, st, gmap)
                                 : launch_grouped<OP_N, 2, 0>(dG, ngroups, blocks, st, gmap);
        else gauss ? launch_grouped<OP_N, 0, 1>(dG, ngroups, blocks, st, gmap)
                   : launch_grouped<OP_N, 0, 0>(dG, ngroups, blocks, st, gmap);
    } else if (opb == OP_C) {
        gauss ? launch_grouped<OP_C, 2, 1>(dG, ngroups, blocks, st, gmap)
              : launch_grouped<OP_C, 2, 0>(dG, ngroups, blocks, st, gmap);
    } else {
        return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

// Heuristic K-split: enough (tile, split) workgroups to fill the chip, at
// least 256 k per split.
template <int OPA, int OPB, int GAUSS>
static void launch_grouped_ab(const ZGroup* dG, int ngroups, int64_t blocks, hipStream_t st,
                              const int* gmap) {
    hipLaunchKernelGGL((zgemm_mfma_kernel<OPA, OPB, 0, GAUSS, 1>), dim3((unsigned)blocks),
                       dim3(256), 0, st, 0, 0, 0, (const double2*)nullptr, (int64_t)0, (int64_t)0,
                       (const double2*)nullptr, (int64_t)0, (int64_t)0, (double2*)nullptr,
                       (int64_t)0, (int64_t)0, make_double2(0, 0), make_double2(0, 0), 1, 0,
                       (double2*)nullptr, dG, ngroups, gmap);
}

int zgemm_grouped_ops(int opa, int opb, int tag, int gauss, const ZGroup* dG, int ngroups,
                      int64_t blocks, hipStream_t st, const int* gmap) {
    if (ngroups <= 0 || blocks <= 0) return 0;
    if (opa == OP_N && (opb == OP_N || opb == OP_C))
        return zgemm_grouped(opb, tag, gauss, dG, ngroups, blocks, st, gmap);
#define RCC_G(a, b)                                                                           \
    if (opa == a && opb == b) {                                                               \
        gauss ? launch_grouped_ab<a, b, 1>(dG, ngroups, blocks, st, gmap)                     \
              : launch_grouped_ab<a, b, 0>(dG, ngroups, blocks, st, gmap);                    \
        return hipGetLastError() == hipSuccess ? 0 : -2;                                       \
    }
    RCC_G(OP_N, OP_T)
    RCC_G(OP_T, OP_N)
    RCC_G(OP_C, OP_N)
#undef RCC_G
    return -1;
}

// The grouped GEMM on the 128 x 32 tile (waves of 64 x 16): for groups
// whose n suits 32-column tiles better than 64-column ones (n = 65 .. 96
// needs two 64-wide tiles, mostly padding, but three 32-wide ones).  Same
// MFMA instruction and k order as the 96 x 64 tile: the same bits.
int64_t zgroup_offsets_tall(ZGroup* g, int ngroups) {
    int64_t off = 0;
    for (int i = 0; i < ngroups; ++i) {
        g[i].tile_off = off;
        const int64_t t =
            (int64_t)((g[i].m + TLS::BM - 1) / TLS::BM) * ((g[i].n + TLS::BN - 1) / TLS::BN);
        off += (t + NXCD - 1) / NXCD * NXCD;
    }
    return off;
}

template <int OPA, int OPB, int GAUSS>
static void launch_grouped_tall(const ZGroup* dG, int ngroups, int64_t blocks, hipStream_t st) {
    hipLaunchKernelGGL((zgemm_mfma_kernel<OPA, OPB, 0, GAUSS, 1, TLS>), dim3((unsigned)blocks),
                       dim3(256), 0, st, 0, 0, 0, (const double2*)nullptr, (int64_t)0, (int64_t)0,
                       (const double2*)nullptr, (int64_t)0, (int64_t)0, (double2*)nullptr,
                       (int64_t)0, (int64_t)0, make_double2(0, 0), make_double2(0, 0), 1, 0,
                       (double2*)nullptr, dG, ngroups, nullptr);
}

int zgemm_grouped_tall(int opa, int opb, int gauss, const ZGroup* dG, int ngroups,
                       int64_t blocks, hipStream_t st) {
    if (ngroups <= 0 || blocks <= 0) return 0;
#define RCC_T(a, b)                                                                           \
    if (opa == a && opb == b) {                                                               \
        gauss ? launch_grouped_tall<a, b, 1>(dG, ngroups, blocks, st)                         \
              : launch_grouped_tall<a, b, 0>(dG, ngroups, blocks, st);                        \
        return hipGetLastError() == hipSuccess ? 0 : -2;                                       \
    }
    RCC_T(OP_N, OP_N)
    RCC_T(OP_N, OP_C)
    RCC_T(OP_N, OP_T)
    RCC_T(OP_T, OP_N)
    RCC_T(OP_C, OP_N)
#undef RCC_T
    return -1;
}

// Streaming grouped product for C (m x n) = alpha A B + beta C with op N N,
// n <= NJ and m large: one row of C per thread, the row of A read once
// (coalesced across the wave along m), B staged in LDS in k-chunks of 32 and
// read as broadcasts, FP64 FMAs on the vector units.  For the products
// whose block has an occupied-spinor dimension (n = o of one irrep): every
// MFMA tile would be mostly padding and the product is bound by reading A.
constexpr int SKR = 256, SKK = 32;

int64_t zgroup_offsets_skinny(ZGroup* g, int ngroups) {
    int64_t off = 0;
    for (int i = 0; i < ngroups; ++i) {
        g[i].tile_off = off;
        off += (g[i].m + SKR - 1) / SKR;
    }
    return off;
}

template <int NJ>
__global__ void __launch_bounds__(256)
zgemm_skinny_grouped_kernel(const ZGroup* __restrict__ G, int ngroups) {
    __shared__ double2 Bs[SKK][NJ];
    const int64_t b = blockIdx.x;
    int lo = 0, hi = ngroups - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (G[mid].tile_off <= b) lo = mid;
        else hi = mid - 1;
    }
    const ZGroup& gg = G[lo];
    const int m = gg.m, n = gg.n, k = gg.k;
    const int64_t row = (b - gg.tile_off) * SKR + threadIdx.x;
    const bool live = row < m;
    const double2* __restrict__ A = gg.A;
    const double2* __restrict__ B = gg.B;
    double ar[NJ], ai[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        ar[j] = 0.0;
        ai[j] = 0.0;
    }
    for (int k0 = 0; k0 < k; k0 += SKK) {
        __syncthreads();
        for (int t = threadIdx.x; t < SKK * NJ; t += SKR) {
            const int kk = t / NJ, j = t - (t / NJ) * NJ;
            Bs[kk][j] = (k0 + kk < k && j < n) ? B[(k0 + kk) + (int64_t)j * gg.ldb]
                                                : make_double2(0.0, 0.0);
        }
        __syncthreads();
        const int kl = min(SKK, k - k0);
        if (live)
            for (int kk = 0; kk < kl; ++kk) {
                const double2 a = A[row + (int64_t)(k0 + kk) * gg.lda];
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    const double2 bb = Bs[kk][j];
                    ar[j] = fma(a.x, bb.x, ar[j]);
                    ar[j] = fma(-a.y, bb.y, ar[j]);
                    ai[j] = fma(a.x, bb.y, ai[j]);
                    ai[j] = fma(a.y, bb.x, ai[j]);
                }
            }
    }
    if (!live) return;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        if (j >= n) break;
        double2* cp = gg.C + row + (int64_t)j * gg.ldc;
        double2 o;
        o.x = gg.alpha.x * ar[j] - gg.alpha.y * ai[j];
        o.y = gg.alpha.x * ai[j] + gg.alpha.y * ar[j];
        if (gg.beta.x != 0.0 || gg.beta.y != 0.0) {
            const double2 c = *cp;
            o.x += gg.beta.x * c.x - gg.beta.y * c.y;
            o.y += gg.beta.x * c.y + gg.beta.y * c.x;
        }
        *cp = o;
    }
}

int zgemm_grouped_skinny(int nj, const ZGroup* dG, int ngroups, int64_t blocks, hipStream_t st) {
    if (ngroups <= 0 || blocks <= 0) return 0;
    if (nj <= 4)
        hipLaunchKernelGGL(zgemm_skinny_grouped_kernel<4>, dim3((unsigned)blocks), dim3(SKR), 0,
                           st, dG, ngroups);
    else if (nj <= 8)
        hipLaunchKernelGGL(zgemm_skinny_grouped_kernel<8>, dim3((unsigned)blocks), dim3(SKR), 0,
                           st, dG, ngroups);
    else if (nj <= 16)
        hipLaunchKernelGGL(zgemm_skinny_grouped_kernel<16>, dim3((unsigned)blocks), dim3(SKR), 0,
                           st, dG, ngroups);
    else
        return -1;
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int choose_split(const GemmArgs& g, int64_t ws_elems) {
    const int64_t tiles = tile_count(g);
    if (tiles >= 512 || g.k < 512) return 1;
    int64_t s = std::min<int64_t>((1024 + tiles - 1) / tiles, (g.k + 255) / 256);
    s = std::min<int64_t>(s, 256);
    const int64_t per = (int64_t)g.m * g.n * g.batch;
    if (per > 0) s = std::min<int64_t>(s, ws_elems / per);
    return (int)std::max<int64_t>(1, s);
}

// (nsplit, kchunk) for a requested split count: k chunks a multiple of ZBK.
static void plan_split(const GemmArgs& g, int s, int& nsplit, int& kchunk) {
    nsplit = 1;
    kchunk = g.k;
    if (s <= 1) return;
    kchunk = (g.k + s - 1) / s;
    kchunk = (kchunk + ZBK - 1) / ZBK * ZBK;
    nsplit = (g.k + kchunk - 1) / kchunk;
    if (nsplit <= 1) {
        nsplit = 1;
        kchunk = g.k;
    }
}

void launch_split_reduce(int m, int n, int nsplit, const double2* W, double2* C, int64_t ldc,
                         double2 alpha, double2 beta, hipStream_t st) {
    const int64_t tot = (int64_t)m * n * split_ways(nsplit);
    const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65535);
    hipLaunchKernelGGL(zgemm_reduce_kernel, dim3(blocks), dim3(256), 0, st, m, n, nsplit, 1, W, C,
                       ldc, (int64_t)0, alpha, beta);
}

static int launch_split(const GemmArgs& g, int nsplit, int kchunk, double2* ws, hipStream_t st) {
    double2* Wm = nsplit > 1 ? ws : nullptr;
#define RCC_L(a, b) if (g.opA == a && g.opB == b) { launch<a, b>(g, nsplit, kchunk, Wm, st); }
    RCC_L(OP_N, OP_N) else RCC_L(OP_N, OP_T) else RCC_L(OP_N, OP_C)
    else RCC_L(OP_T, OP_N) else RCC_L(OP_T, OP_T) else RCC_L(OP_T, OP_C)
    else RCC_L(OP_C, OP_N) else RCC_L(OP_C, OP_T) else RCC_L(OP_C, OP_C)
    else RCC_L(OP_R, OP_N) else RCC_L(OP_N, OP_R)
    else return -1;
#undef RCC_L
    if (nsplit > 1) {
        const int64_t tot = (int64_t)g.m * g.n * g.batch * split_ways(nsplit);
        const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65535);
        hipLaunchKernelGGL(zgemm_reduce_kernel, dim3(blocks), dim3(256), 0, st, g.m, g.n, nsplit,
                           g.batch, Wm, g.C, g.ldc, g.sC, g.alpha, g.beta);
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ---- K-split autotuner
// A product whose output tiles cannot fill the chip (fewer than 512 tile x
// batch workgroups), or whose tiles leave the last round of workgroups
// poorly filled (less than 95 % of the workgroups resident at once: two
// per CU), and whose k is long, is split over k; how finely is a trade of
// per-workgroup latency (k stages in a chain), of the round fill, and of
// the partial sums written and reduced.  The first time a shape (ops, tag,
// complex form, m, n, k, batch) is met outside graph capture, every split
// with k chunks of 64 .. 4096, 2 .. 8 equal chunks, the splits whose
// workgroups fill 1 .. 8 whole rounds of the resident slots, and no split, is
// timed on the stream with HIP events.  The trial products write a scratch
// C of the product's size, allocated for the trial and freed after it, so
// the live C is never touched; the fastest split is kept for the process.
// RCC_AUTOTUNE=0 in the environment (or zgemm_autotune(0)) falls back to
// choose_split.
namespace {
struct TuneKey {
    int opa, opb, tag, gauss, m, n, k, batch;
    bool operator<(const TuneKey& o) const {
        const int a[8] = {opa, opb, tag, gauss, m, n, k, batch};
        const int b[8] = {o.opa, o.opb, o.tag, o.gauss, o.m, o.n, o.k, o.batch};
        return std::lexicographical_compare(a, a + 8, b, b + 8);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tuned;
int g_autotune = -1;     // -1: read RCC_AUTOTUNE on first use

bool autotune_on() {
    if (g_autotune < 0) {
        const char* e = std::getenv("RCC_AUTOTUNE");
        g_autotune = (e && e[0] == '0') ? 0 : 1;
    }
    return g_autotune == 1;
}

bool capturing(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess) return true;
    return cs != hipStreamCaptureStatusNone;
}

// fraction of the workgroup slots the tiles keep busy over their rounds
double round_fill(int64_t tiles) {
    const int64_t slots = resident_slots();
    return (double)tiles / (double)(((tiles + slots - 1) / slots) * slots);
}

// Times the candidate splits; returns the fastest split count.
int tune_split(const GemmArgs& g, double2* ws, int64_t ws_elems, hipStream_t st, int fallback) {
    const int64_t tiles = tile_count(g);
    const int64_t per = (int64_t)g.m * g.n * g.batch;
    std::vector<int> cand{1};
    auto add = [&](int64_t s) {
        if (s < 2 || s > 4096 || per * s > ws_elems || tiles * s > 65536) return;
        int ns, kc;
        plan_split(g, (int)s, ns, kc);
        if (ns > 1 && std::find(cand.begin(), cand.end(), ns) == cand.end()) cand.push_back(ns);
    };
    for (int c = 64; c <= 4096; c *= 2) add((g.k + c - 1) / c);
    for (int e = 2; e <= 8; ++e) add(e);
    // splits whose workgroups fill whole rounds of the resident slots
    const int64_t slots = resident_slots();
    for (int r = 1; r <= 8; ++r) add(r * slots / std::max<int64_t>(tiles, 1));
    if (cand.size() == 1) return fallback;
    // trial products on a zeroed scratch C (same shape, tightly packed)
    GemmArgs t = g;
    double2* ct = nullptr;
    if (hipMalloc((void**)&ct, (size_t)per * sizeof(double2)) != hipSuccess) {
        (void)hipGetLastError();
        return fallback;
    }
    (void)hipMemsetAsync(ct, 0, (size_t)per * sizeof(double2), st);
    t.C = ct;
    t.ldc = g.m;
    t.sC = (int64_t)g.m * g.n;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(ct);
        return fallback;
    }
    int best = fallback;
    float best_ms = 1e30f;
    std::vector<std::pair<int, float>> timed;
    for (int s : cand) {
        int ns, kc;
        plan_split(t, s, ns, kc);
        float ms_min = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            (void)hipEventRecord(e0, st);
            if (launch_split(t, ns, kc, ws, st) != 0) { ms_min = 1e30f; break; }
            (void)hipEventRecord(e1, st);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0) ms_min = std::min(ms_min, ms);   // the first run warms caches
        }
        timed.push_back({ns, ms_min});
        if (ms_min < best_ms) {
            best_ms = ms_min;
            best = ns;
        }
    }
    // the tagged products (ladder, rings): a finer split within 1 % of the
    // fastest is kept instead, since its workgroups cover shorter k ranges
    // and refetch fewer operand panels (the 30o/300v ladder: 75 GB in four
    // parts against 92 GB unsplit, in the same time)
    if (g.tag != 0)
        for (const auto& c : timed)
            if (c.second <= best_ms * 1.01f && c.first > best) best = c.first;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(st);
    (void)hipFree(ct);
    return best;
}
}  // namespace

void zgemm_autotune(int enable) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_autotune = enable ? 1 : 0;
}

bool zgemm_autotune_enabled() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return autotune_on();
}

int64_t zgemm_tuned_shapes() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return (int64_t)g_tuned.size();
}

int zgemm(const GemmArgs& g0, double2* ws, int64_t ws_elems, hipStream_t st) {
    GemmArgs g = g0;
    if (g_record) {
        // recorded for a batched replay: the split the tuner chose for this
        // shape (no tuning while recording), else the heuristic
        Rec r;
        r.kind = RK_GEMM;
        r.g = g;
        r.ws = ws;
        int s = ws ? choose_split(g, ws_elems) : 1;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tuned.find(TuneKey{g.opA, g.opB, g.tag, g.gauss, g.m, g.n, g.k, g.batch});
            if (ws && it != g_tuned.end()) s = it->second;
        }
        plan_split(g, s, r.nsplit, r.kchunk);
        g_record->push_back(std::move(r));
        return 0;
    }
    if (g.m <= 0 || g.n <= 0 || g.batch <= 0) return 0;
    if (g.k <= 0) {
        const int64_t tot = (int64_t)g.m * g.n * g.batch;
        const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65535);
        hipLaunchKernelGGL(zscale_kernel, dim3(blocks), dim3(256), 0, st, g.m, g.n, g.batch,
                           g.C, g.ldc, g.sC, g.beta);
        return 0;
    }
    int s = ws ? choose_split(g, ws_elems) : 1;
    const int64_t tiles = tile_count(g);
    const bool few = tiles < 512 && g.k >= 256;
    const bool ragged = g.k >= 2048 && round_fill(tiles) < 0.95;
    if (ws && (few || ragged)) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        if (autotune_on()) {
            const TuneKey key{g.opA, g.opB, g.tag, g.gauss, g.m, g.n, g.k, g.batch};
            auto it = g_tuned.find(key);
            if (it != g_tuned.end()) s = it->second;
            else if (!capturing(st)) s = g_tuned[key] = tune_split(g, ws, ws_elems, st, s);
        }
    }
    int nsplit, kchunk;
    plan_split(g, s, nsplit, kchunk);
    return launch_split(g, nsplit, kchunk, ws, st);
}

// ---- allocation-free products (rcc_zgemm_device_noalloc)
// The K-split partial sums of these products live in this file-scope array
// instead of a workspace the library allocates: 32 MiB (2 Mi complex
// elements) on every GPU the library runs on, held until the process ends.
// Host code never uses the name `scratch` as an address (on the host it is
// only a stand-in); zgemm_scratch_init takes the array's device address once
// per device with that device current, and every launch uses that address.
// All products on one GPU share the array, and it keeps what the last call
// left: calls that overlap on one GPU race.  Nothing needs clearing, since
// every split writes every element of its partial-sum slice before the
// reduction reads it.
namespace {
constexpr int64_t kScratchElems = (int64_t)(32 << 20) / (int64_t)sizeof(double2);
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double2* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) double2 scratch[kScratchElems];

int64_t zgemm_scratch_elems() { return kScratchElems; }

int zgemm_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return 0;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return -1;
    }
    g_scratch_ptr[device] = (double2*)ptr;
    return 0;
}

int zgemm_noalloc(const GemmArgs& g, int device, int ksplit, hipStream_t st, std::string& err) {
    double2* ws = nullptr;
    if (device >= 0 && device < kMaxDevices) {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        ws = g_scratch_ptr[device];
    }
    if (!ws) {
        err = "no scratch address for this device (the handle was not created on it)";
        return -1;
    }
    if (g.m <= 0 || g.n <= 0 || g.batch <= 0) return 0;
    if (g.k <= 0) return zgemm(g, nullptr, 0, st);      // C = beta C, no partial sums
    const int s = ksplit > 0 ? ksplit : choose_split(g, kScratchElems);
    int nsplit, kchunk;
    plan_split(g, s, nsplit, kchunk);
    const int64_t need = nsplit > 1 ? (int64_t)nsplit * g.m * g.n * g.batch : 0;
    if (need > kScratchElems) {
        err = "the K split needs " + std::to_string(need * (int64_t)sizeof(double2)) +
              " bytes of partial sums; the scratch holds " +
              std::to_string(kScratchElems * (int64_t)sizeof(double2)) + " bytes";
        return -1;
    }
    const int rc = launch_split(g, nsplit, kchunk, ws, st);
    if (rc != 0) err = "zgemm launch failed (" + std::to_string(rc) + ")";
    return rc;
}

// Layout probe: one v_mfma_f64_16x16x4_f64 with A[i][k] = i + 16 k and
// B[k][j] = (k == j % 4) ? 1 : 0... the host checks D against the assumed
// lane mapping.  out[64 * 4]: lane-major D registers.
__global__ void mfma_probe_kernel(double* out) {
    const int l = threadIdx.x;
    const int row = l & 15, kk = l >> 4;
    const double a = (double)(row + 16 * kk);
    const double b = (double)((kk + 1) * 1000 + (l & 15));
    v4d c = (v4d){0.0, 0.0, 0.0, 0.0};
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    for (int r = 0; r < 4; ++r) out[l * 4 + r] = c[r];
}

// Throughput probe: every wave runs `iters` rounds of CH independent
// v_mfma_f64_16x16x4_f64 chains on register data (no memory traffic in the
// loop); the host times it to obtain the achievable FP64 matrix-core rate.
template <int CH>
__global__ void __launch_bounds__(256) mfma_rate_kernel(int iters, double seed, double* out) {
    v4d acc[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = (v4d){0.0, 0.0, 0.0, 0.0};
    double a = seed * (threadIdx.x + 1), b = seed * 0.5;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int c = 0; c < CH; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
    }
    double t = 0.0;
#pragma unroll
    for (int c = 0; c < CH; ++c) t += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
    if (t == 12345.678) out[threadIdx.x] = t;   // keeps the chains live, never true for the probe
}

int mfma_rate(int chains, int iters, int nblocks, double* out, hipStream_t st) {
    switch (chains) {
        case 1: hipLaunchKernelGGL(mfma_rate_kernel<1>, dim3(nblocks), dim3(256), 0, st, iters, 1e-30, out); break;
        case 2: hipLaunchKernelGGL(mfma_rate_kernel<2>, dim3(nblocks), dim3(256), 0, st, iters, 1e-30, out); break;
        case 4: hipLaunchKernelGGL(mfma_rate_kernel<4>, dim3(nblocks), dim3(256), 0, st, iters, 1e-30, out); break;
        case 8: hipLaunchKernelGGL(mfma_rate_kernel<8>, dim3(nblocks), dim3(256), 0, st, iters, 1e-30, out); break;
        case 12: hipLaunchKernelGGL(mfma_rate_kernel<12>, dim3(nblocks), dim3(256), 0, st, iters, 1e-30, out); break;
        default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mfma_probe(double* dout, hipStream_t st) {
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, st, dout);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace rcc
