// Split-INT8 complex GEMM with exact integer accumulation (Ozaki-style).
//
// Representation.  An operand X (R rows x K, complex) is carried as three real
// variants, Re X, Im X and Re X + Im X (the Gauss three-product form).  For
// each variant, each row and each chunk of 8192 k there is a power-of-two
// scale 2^e with e the frexp exponent of the chunk's largest magnitude, so the
// scaled values x = v 2^-e lie in (-1, 1).  x is written as S signed digits
//     x ~ 2^-6 (d0 + d1 2^-7 + d2 2^-14 + ...),   d_i in [-64, 64],
// by successive rounding (d0 = rint(64 x), then rint(128 r) of each residual),
// so the representation error is at most 2^-(7S) of the chunk maximum.
//
// Product.  A digit product d_i d'_j is an exact integer below 2^12, and
// v_mfma_i32_16x16x64_i8 sums them exactly in int32.  The pairs with
// i + j = o (o < S) share one accumulator per variant, so a 16 x 16 output
// block holds 3 S int32 accumulators; nothing is rounded until a chunk is
// complete, when each accumulator is scaled by 2^(-12-7o) 2^(eA+eB) and
// combined in FP64: Re = P0 - P1, Im = P2 - P0 - P1.
//
// Layout in HBM.  Digits are stored fragment-major: fragment (row block of 16,
// k-step of 64, variant, digit) is 1 KiB, lane l holding the 16 bytes of row
// 16 rb + l % 16, k = 64 ks + 16 (l / 16) + 0..15, which is exactly the lane's
// MFMA operand.  One wave loads a fragment with one coalesced 16-byte load per
// lane, writes it to LDS contiguously and reads it back conflict-free.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include "rcc_internal.h"

namespace rcc {

typedef int i32x4 __attribute__((ext_vector_type(4)));

constexpr int I8_KSTEP = 64;                       // k per MFMA step / fragment
constexpr int I8_KC = 8192;                        // k per scale chunk
constexpr int I8_KS_PER_CH = I8_KC / I8_KSTEP;     // 128

__device__ __forceinline__ int chunk_exp(unsigned long long bits) {
    const double mx = __longlong_as_double((long long)bits);
    int e = 0;
    if (mx > 0.0) (void)frexp(mx, &e);
    return e;
}

// ---------------------------------------------------------------- tile load
// 64 rows x 64 k of X (SRC 0: X[r + k ld], SRC 1: X[k + r ld]) into LDS,
// zero outside R x K, conjugated on request.
template <int SRC>
__device__ __forceinline__ void load_tile(double2 (*t)[65], const double2* __restrict__ X,
                                          int64_t ld, int R, int K, int conj, int r0, int k0) {
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        int r, k;
        if (SRC == 0) { r = e % 64; k = e / 64; }
        else { k = e % 64; r = e / 64; }
        const int gr = r0 + r, gk = k0 + k;
        double2 v = make_double2(0.0, 0.0);
        if (gr < R && gk < K) v = SRC == 0 ? X[gr + (int64_t)gk * ld] : X[gk + (int64_t)gr * ld];
        if (conj) v.y = -v.y;
        t[r][k] = v;
    }
}

__device__ __forceinline__ double variant(double2 v, int which) {
    return which == 0 ? v.x : (which == 1 ? v.y : v.x + v.y);
}

// mx[(row nch + chunk) 3 + variant] = bit pattern of the largest |value|
__global__ void __launch_bounds__(256) i8_clear_kernel(unsigned long long* mx, int64_t n) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        mx[i] = 0ull;
}

template <int SRC>
__global__ void __launch_bounds__(256) i8_max_kernel(const double2* __restrict__ X, int64_t ld,
                                                     int R, int K, int conj, int nch,
                                                     unsigned long long* __restrict__ mx) {
    __shared__ double2 t[64][65];
    const int ks = blockIdx.x, r0 = blockIdx.y * 64;
    load_tile<SRC>(t, X, ld, R, K, conj, r0, ks * I8_KSTEP);
    __syncthreads();
    if (threadIdx.x < 192) {
        const int r = threadIdx.x % 64, w = threadIdx.x / 64;
        double m = 0.0;
        for (int k = 0; k < 64; ++k) m = fmax(m, fabs(variant(t[r][k], w)));
        if (m > 0.0)
            atomicMax(mx + ((int64_t)(r0 + r) * nch + ks / I8_KS_PER_CH) * 3 + w,
                      (unsigned long long)__double_as_longlong(m));
    }
}

template <int SRC, int S>
__global__ void __launch_bounds__(256) i8_digits_kernel(const double2* __restrict__ X, int64_t ld,
                                                        int R, int K, int conj, int nks, int nch,
                                                        const unsigned long long* __restrict__ mx,
                                                        i32x4* __restrict__ P) {
    __shared__ double2 t[64][65];
    const int ks = blockIdx.x, r0 = blockIdx.y * 64;
    load_tile<SRC>(t, X, ld, R, K, conj, r0, ks * I8_KSTEP);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = wave * 16 + (lane & 15), g = lane >> 4;
    const int rb = (r0 >> 4) + wave;
    const int64_t mrow = ((int64_t)(r0 + r) * nch + ks / I8_KS_PER_CH) * 3;
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        const int e = chunk_exp(mx[mrow + w]);
        unsigned int words[S][4];
#pragma unroll
        for (int d = 0; d < S; ++d)
#pragma unroll
            for (int q = 0; q < 4; ++q) words[d][q] = 0u;
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            double x = ldexp(variant(t[r][g * 16 + b], w), -e) * 64.0;
#pragma unroll
            for (int d = 0; d < S; ++d) {
                const double q = rint(x);
                x = (x - q) * 128.0;
                words[d][b >> 2] |= ((unsigned int)(int)q & 0xffu) << (8 * (b & 3));
            }
        }
#pragma unroll
        for (int d = 0; d < S; ++d) {
            const int64_t f = (((int64_t)rb * nks + ks) * 3 + w) * S + d;
            P[f * 64 + lane] = (i32x4){(int)words[d][0], (int)words[d][1], (int)words[d][2],
                                       (int)words[d][3]};
        }
    }
}

// ---------------------------------------------------------------- GEMM
// C(m x n) = alpha A B^T + beta C.  One 256-thread workgroup per 64 x 64 tile,
// waves 2 x 2 of 32 x 32 (2 x 2 MFMA blocks).  A stage is one (k-step,
// variant): the S digit fragments of the tile's 4 A row blocks and 4 B row
// blocks (8 S KiB), double-buffered in LDS with one barrier per stage.
// Stages run chunk by chunk, and within a chunk variant by variant, so only
// one variant's S int32 accumulators are live; after each (chunk, variant)
// they are scaled and added into C (FP64).
struct I8Tile {
    int m, n, nks, nch, tm, tn, wm, wn, lane, wave;
    int row0, col0;   // first row / column of this wave's 32 x 32 block
    const unsigned long long* MA;
    const unsigned long long* MB;
    double2* C;
    int64_t ldc;
    double2 alpha, beta;
};

// (chunk, variant) complete: P = sum_o acc_o 2^(-12-7o) 2^(eA+eB) enters C as
// alpha (P, -P), alpha (-P, -P) or alpha (0, P) for variants Re, Im, Re+Im
// (Re = P0 - P1, Im = P2 - P0 - P1); the first flush also applies beta.
// Each element of C has one owning lane.  Clears the accumulators.
template <int S>
__device__ __forceinline__ void i8_flush(const I8Tile& t, int c, int v, i32x4 (&acc)[S][2][2]) {
    const double sre = v == 0 ? 1.0 : (v == 1 ? -1.0 : 0.0);
    const double sim = v == 2 ? 1.0 : -1.0;
    const bool first = c == 0 && v == 0;
    int eb[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int col = t.col0 + j * 16 + (t.lane & 15);
        eb[j] = chunk_exp(t.MB[((int64_t)col * t.nch + c) * 3 + v]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = t.row0 + i * 16 + 4 * (t.lane >> 4) + r;
            const int ea = chunk_exp(t.MA[((int64_t)row * t.nch + c) * 3 + v]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                double p = 0.0;
#pragma unroll
                for (int o = S - 1; o >= 0; --o) p += ldexp((double)acc[o][i][j][r], -12 - 7 * o);
                p = ldexp(p, ea + eb[j]);
                const double xr = sre * p, xi = sim * p;
                const int col = t.col0 + j * 16 + (t.lane & 15);
                if (row < t.m && col < t.n) {
                    double2* cp = t.C + row + (int64_t)col * t.ldc;
                    double2 o;
                    o.x = t.alpha.x * xr - t.alpha.y * xi;
                    o.y = t.alpha.x * xi + t.alpha.y * xr;
                    if (!first) {
                        const double2 cv = *cp;
                        o.x += cv.x;
                        o.y += cv.y;
                    } else if (t.beta.x != 0.0 || t.beta.y != 0.0) {
                        const double2 cv = *cp;
                        o.x += t.beta.x * cv.x - t.beta.y * cv.y;
                        o.y += t.beta.x * cv.y + t.beta.y * cv.x;
                    }
                    *cp = o;
                }
            }
        }
#pragma unroll
    for (int o = 0; o < S; ++o)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[o][i][j] = (i32x4){0, 0, 0, 0};
}

// s_waitcnt vmcnt(N) only (expcnt and lgkmcnt at their no-wait maxima).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

typedef __attribute__((address_space(3))) void* lds_ptr;

// Stage (ks, v) of a (chunk, variant)-ordered sequence: chunk c, variant v,
// k-step ks; `last` when the next stage starts another (chunk, variant).
struct I8Stage {
    int c, v, ks;
    bool last;
};

__device__ __forceinline__ I8Stage i8_stage_of(int st, int nks) {
    const int full = nks / I8_KS_PER_CH, rem = nks % I8_KS_PER_CH;
    I8Stage s;
    int r, len;
    if (st < full * 3 * I8_KS_PER_CH) {
        s.c = st / (3 * I8_KS_PER_CH);
        r = st % (3 * I8_KS_PER_CH);
        len = I8_KS_PER_CH;
    } else {
        s.c = full;
        r = st - full * 3 * I8_KS_PER_CH;
        len = rem;
    }
    s.v = r / len;
    s.ks = s.c * I8_KS_PER_CH + r % len;
    s.last = r % len == len - 1;
    return s;
}

// Asynchronous copy of stage st into LDS buffer `dst` (global -> LDS DMA, no
// registers): wave w copies the S digit fragments of A row block w and, if
// w < 4, of B row block w; lane l writes bytes 16 l .. 16 l + 15 of each
// 1 KiB fragment, exactly the fragment-major layout of the planes.
template <int S, int NA>
__device__ __forceinline__ void i8_copy(i32x4 (*dst)[64], const i32x4* ga, const i32x4* gb,
                                        int wave, int st, int nks) {
    const I8Stage g = i8_stage_of(st, nks);
    const int64_t off = ((int64_t)g.ks * 3 + g.v) * S * 64;
#pragma unroll
    for (int d = 0; d < S; ++d)
        __builtin_amdgcn_global_load_lds((const void*)(ga + off + d * 64),
                                         (lds_ptr)&dst[wave * S + d][0], 16, 0, 0);
    if (wave < 4) {
#pragma unroll
        for (int d = 0; d < S; ++d)
            __builtin_amdgcn_global_load_lds((const void*)(gb + off + d * 64),
                                             (lds_ptr)&dst[(NA + wave) * S + d][0], 16, 0, 0);
    }
}

// C(m x n) = alpha A B^T + beta C.  One workgroup of 2 WM waves per
// (32 WM) x 64 tile, waves WM x 2 of 32 x 32 (2 x 2 MFMA blocks).  A stage is
// one (k-step, variant): the S digit fragments of the tile's 2 WM A and 4 B
// row blocks.  Three LDS buffers: stage st + 2 is copied by DMA while st is
// multiplied, and a wave waits only for its own copies of stage st before
// the stage barrier.  Stages run chunk by chunk, and within a chunk variant
// by variant, so only one variant's S int32 accumulators are live; after each
// (chunk, variant) they are scaled and added into C (FP64).
//
// Tile order: each XCD (workgroups dealt round-robin) takes a contiguous
// range of tile numbers; gm > 1 numbers the tiles in bands of gm tile rows,
// column by column inside a band, so the workgroups an XCD runs together
// cover a gm x (32 / gm) block of tiles and share A and B panels through that
// XCD's L2 (gm = 1: row-major, one A panel and 32 B panels).
template <int S, int WM>
__global__ void __launch_bounds__(128 * WM) zgemm_i8_kernel(
    int m, int n, int nks, int nch, const i32x4* __restrict__ PA,
    const unsigned long long* __restrict__ MA, const i32x4* __restrict__ PB,
    const unsigned long long* __restrict__ MB, double2* __restrict__ C, int64_t ldc,
    double2 alpha, double2 beta, int gm) {
    constexpr int NA = 2 * WM;
    __shared__ i32x4 lds[3][(NA + 4) * S][64];
    const int mt = (m + 32 * WM - 1) / (32 * WM), nt = (n + 63) / 64;
    const int T = mt * nt;
    const int b = blockIdx.x;
    const int xcd = b % 8, idx = b / 8;
    const int q = T / 8, rr = T % 8;
    const int start = xcd < rr ? xcd * (q + 1) : rr * (q + 1) + (xcd - rr) * q;
    const int tile = start + idx;
    I8Tile t;
    t.m = m;
    t.n = n;
    t.nks = nks;
    t.nch = nch;
    {
        const int band = gm * nt;                 // tiles per band of gm tile rows
        const int first = (tile / band) * gm;
        const int rows = min(mt - first, gm);
        const int r = tile % band;
        t.tm = first + r % rows;
        t.tn = r / rows;
    }
    t.lane = threadIdx.x & 63;
    t.wave = threadIdx.x >> 6;
    t.wm = t.wave % WM;
    t.wn = t.wave / WM;
    t.row0 = t.tm * 32 * WM + t.wm * 32;
    t.col0 = t.tn * 64 + t.wn * 32;
    t.MA = MA;
    t.MB = MB;
    t.C = C;
    t.ldc = ldc;
    t.alpha = alpha;
    t.beta = beta;
    const int64_t rbs = (int64_t)nks * 3 * S * 64;   // i32x4 per row block
    const i32x4* ga = PA + (NA * t.tm + t.wave) * rbs + t.lane;
    const i32x4* gb = PB + (4 * t.tn + (t.wave & 3)) * rbs + t.lane;
    const bool copy_b = t.wave < 4;

    i32x4 acc[S][2][2];
#pragma unroll
    for (int o = 0; o < S; ++o)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[o][i][j] = (i32x4){0, 0, 0, 0};

    const int nst = 3 * nks;
    i8_copy<S, NA>(lds[0], ga, gb, t.wave, 0, nks);
    if (nst > 1) i8_copy<S, NA>(lds[1], ga, gb, t.wave, 1, nks);
    for (int st = 0; st < nst; ++st) {
        // this wave's copies of stage st have landed (those of st + 1 may not)
        if (st + 1 < nst) {
            if (copy_b) wait_vmcnt<2 * S>();
            else wait_vmcnt<S>();
        } else {
            wait_vmcnt<0>();
        }
        __builtin_amdgcn_s_barrier();
        // every wave is past stage st - 1, so its buffer can be refilled
        if (st + 2 < nst) i8_copy<S, NA>(lds[(st + 2) % 3], ga, gb, t.wave, st + 2, nks);
        const int cb = st % 3;
        i32x4 a[2][S], bl[2][S];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int d = 0; d < S; ++d) {
                a[i][d] = lds[cb][(2 * t.wm + i) * S + d][t.lane];
                bl[i][d] = lds[cb][(NA + 2 * t.wn + i) * S + d][t.lane];
            }
#pragma unroll
        for (int da = 0; da < S; ++da)
#pragma unroll
            for (int db = 0; db + da < S; ++db)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[da + db][i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                            a[i][da], bl[j][db], acc[da + db][i][j], 0, 0, 0);
        const I8Stage g = i8_stage_of(st, nks);
        if (g.last) i8_flush<S>(t, g.c, g.v, acc);
    }
}

// ---------------------------------------------------------------- host side
void i8_layout(I8Operand& S, int rows, int k, int digits) {
    S.rows = rows;
    S.k = k;
    S.digits = digits;
    S.rows_pad = (rows + 383) / 384 * 384;   // whole 64-, 96- and 128-row GEMM tiles
    S.nks = (k + I8_KSTEP - 1) / I8_KSTEP;
    S.nch = (S.nks + I8_KS_PER_CH - 1) / I8_KS_PER_CH;
    S.plane_bytes = (int64_t)(S.rows_pad / 16) * S.nks * 3 * digits * 1024;
    S.mx_elems = (int64_t)S.rows_pad * S.nch * 3;
}

template <int SRC, int D>
static void launch_digits(const I8Operand& S, const double2* X, int64_t ld, int conj,
                          hipStream_t st) {
    dim3 g((unsigned)S.nks, (unsigned)(S.rows_pad / 64));
    hipLaunchKernelGGL((i8_digits_kernel<SRC, D>), g, dim3(256), 0, st, X, ld, S.rows, S.k, conj,
                       S.nks, S.nch, S.mx, (i32x4*)S.planes);
}

int i8_split(const I8Operand& S, const double2* X, int64_t ld, bool k_contiguous, bool conj,
             hipStream_t st) {
    if (S.rows <= 0 || S.k <= 0) return 0;
    if (S.digits < 4 || S.digits > 5) return -1;
    const int cj = conj ? 1 : 0;
    const unsigned nb = (unsigned)std::min<int64_t>((S.mx_elems + 255) / 256, 4096);
    hipLaunchKernelGGL(i8_clear_kernel, dim3(nb), dim3(256), 0, st, S.mx, S.mx_elems);
    dim3 g((unsigned)S.nks, (unsigned)(S.rows_pad / 64));
    if (k_contiguous) {
        hipLaunchKernelGGL(i8_max_kernel<1>, g, dim3(256), 0, st, X, ld, S.rows, S.k, cj, S.nch,
                           S.mx);
        if (S.digits == 4) launch_digits<1, 4>(S, X, ld, cj, st);
        else launch_digits<1, 5>(S, X, ld, cj, st);
    } else {
        hipLaunchKernelGGL(i8_max_kernel<0>, g, dim3(256), 0, st, X, ld, S.rows, S.k, cj, S.nch,
                           S.mx);
        if (S.digits == 4) launch_digits<0, 4>(S, X, ld, cj, st);
        else launch_digits<0, 5>(S, X, ld, cj, st);
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <int S, int WM>
static void launch_i8(const I8Operand& A, const I8Operand& B, double2* C, int64_t ldc,
                      double2 alpha, double2 beta, int gm, hipStream_t st) {
    // must match the kernel's tile count; planes cover whole 128-row tiles
    const int64_t tiles = (int64_t)((A.rows + 32 * WM - 1) / (32 * WM)) * ((B.rows + 63) / 64);
    hipLaunchKernelGGL((zgemm_i8_kernel<S, WM>), dim3((unsigned)tiles), dim3(128 * WM), 0, st,
                       A.rows, B.rows, A.nks, A.nch, (const i32x4*)A.planes, A.mx,
                       (const i32x4*)B.planes, B.mx, C, ldc, alpha, beta, std::max(gm, 1));
}

// Tile height WM (tile = 32 WM x 64): three LDS buffers of (2 WM + 4) S KiB,
// 144 KiB for (4, 4), 150 KiB for (5, 3); rows_pad (a multiple of 384) is a
// multiple of every 32 WM offered.
static int launch_wm(int S, int wm, int gm, const I8Operand& A, const I8Operand& B,
                     double2* C, int64_t ldc, double2 alpha, double2 beta, hipStream_t st) {
    if (wm == 8) return launch_i8_wide(S, gm, A, B, C, ldc, alpha, beta, st);
    if (S == 4) {
        if (wm == 2) launch_i8<4, 2>(A, B, C, ldc, alpha, beta, gm, st);
        else if (wm == 3) launch_i8<4, 3>(A, B, C, ldc, alpha, beta, gm, st);
        else launch_i8<4, 4>(A, B, C, ldc, alpha, beta, gm, st);
    } else if (S == 5) {
        if (wm == 2) launch_i8<5, 2>(A, B, C, ldc, alpha, beta, gm, st);
        else launch_i8<5, 3>(A, B, C, ldc, alpha, beta, gm, st);
    } else {
        return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

// Tile autotuner: the first product of a shape (digits, m, n, k steps) met
// outside graph capture times every tile height WM (2, 3, 4 and the 128 x 128
// tile of rcc_split_i8w.hip, "8", for 4 digits; 2, 3 for 5) with tile-row
// bands gm of 1, 4 and 8, in place with alpha = 0,
// beta = 1 (C unchanged: the first flush writes 1 C + 0, later flushes add
// 0), and keeps the fastest for the process.  RCC_I8_WM fixes the height
// (band 1); with the autotuner off (RCC_AUTOTUNE=0, zgemm_autotune(0)) it is
// 4 / 3 with band 1.  The choice does not change the result: every tile sums
// the same products in the same order.
struct I8Choice {
    int wm, gm;
};
static std::mutex i8_tune_mu;
static std::map<std::tuple<int, int, int, int>, I8Choice> i8_tuned;

static I8Choice i8_choose(const I8Operand& A, const I8Operand& B, double2* C, int64_t ldc,
                          hipStream_t st) {
    const int S = A.digits;
    const I8Choice dflt{S == 4 ? 4 : 3, 1};
    const char* e = getenv("RCC_I8_WM");
    if (e) {
        const int w = atoi(e);
        return (w == 2 || w == 3 || ((w == 4 || w == 8) && S == 4)) ? I8Choice{w, 1} : dflt;
    }
    if (!zgemm_autotune_enabled()) return dflt;
    std::lock_guard<std::mutex> lk(i8_tune_mu);
    const auto key = std::make_tuple(S, A.rows, B.rows, A.nks);
    auto it = i8_tuned.find(key);
    if (it != i8_tuned.end()) return it->second;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
        return dflt;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return dflt;
    I8Choice best = dflt;
    float best_ms = 1e30f;
    const int wms4[] = {2, 3, 4, 8}, wms5[] = {2, 3};
    const int* wms = S == 4 ? wms4 : wms5;
    const int nw = S == 4 ? 4 : 2;
    const int gms[] = {1, 4, 8};
    for (int c = 0; c < nw; ++c)
        for (int g : gms) {
            float ms_min = 1e30f;
            for (int rep = 0; rep < 2; ++rep) {
                (void)hipEventRecord(e0, st);
                if (launch_wm(S, wms[c], g, A, B, C, ldc, make_double2(0, 0), make_double2(1, 0),
                              st))
                    break;
                (void)hipEventRecord(e1, st);
                (void)hipEventSynchronize(e1);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (rep > 0) ms_min = std::min(ms_min, ms);
            }
            if (ms_min < best_ms) {
                best_ms = ms_min;
                best = I8Choice{wms[c], g};
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    i8_tuned[key] = best;
    return best;
}

int64_t i8_tuned_shapes() {
    std::lock_guard<std::mutex> lk(i8_tune_mu);
    return (int64_t)i8_tuned.size();
}

int zgemm_i8(const I8Operand& A, const I8Operand& B, double2* C, int64_t ldc, double2 alpha,
             double2 beta, hipStream_t st) {
    if (A.nks != B.nks || A.digits != B.digits) return -1;
    const int m = A.rows, n = B.rows;
    if (m <= 0 || n <= 0) return 0;
    if (A.nks == 0) return -1;
    if (A.digits != 4 && A.digits != 5) return -1;
    const I8Choice ch = i8_choose(A, B, C, ldc, st);
    return launch_wm(A.digits, ch.wm, ch.gm, A, B, C, ldc, alpha, beta, st);
}

}  // namespace rcc
