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
struct I8Stage {
    int c, v, ks;
};

__device__ __forceinline__ bool i8_next(I8Stage& s, int nks) {
    const int ks0 = s.c * I8_KS_PER_CH, ks1 = min(nks, ks0 + I8_KS_PER_CH);
    if (s.ks + 1 < ks1) { ++s.ks; return true; }
    if (s.v < 2) { ++s.v; s.ks = ks0; return true; }
    if (ks1 < nks) { ++s.c; s.v = 0; s.ks = ks1; return true; }
    return false;
}

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

// One stage: issue the global loads of the next stage into registers,
// multiply out of LDS buffer `buf`, write the next stage into the other
// buffer, flush when a (chunk, variant) ends, barrier.  Returns false after
// the last stage.  The tile has NA = 2 WM A row blocks and 4 B row blocks;
// wave w copies A row block w and, if w < 4, B row block w.
template <int S, int WM>
__device__ __forceinline__ bool i8_stage(const I8Tile& t, i32x4 (*lds)[(2 * WM + 4) * S][64],
                                         const i32x4* ga, const i32x4* gb, I8Stage& cur,
                                         int& buf, i32x4 (&acc)[S][2][2]) {
    constexpr int NA = 2 * WM;
    I8Stage n1 = cur;
    const bool has1 = i8_next(n1, t.nks);
    const bool copy_b = t.wave < 4;
    i32x4 sa[S], sb[S];
    if (has1) {
        const int64_t off = ((int64_t)n1.ks * 3 + n1.v) * S * 64;
#pragma unroll
        for (int d = 0; d < S; ++d) {
            sa[d] = ga[off + d * 64];
            if (copy_b) sb[d] = gb[off + d * 64];
        }
    }
    {
        i32x4 a[2][S], bl[2][S];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int d = 0; d < S; ++d) {
                a[i][d] = lds[buf][(2 * t.wm + i) * S + d][t.lane];
                bl[i][d] = lds[buf][(NA + 2 * t.wn + i) * S + d][t.lane];
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
    }
    if (has1) {
#pragma unroll
        for (int d = 0; d < S; ++d) {
            lds[buf ^ 1][t.wave * S + d][t.lane] = sa[d];
            if (copy_b) lds[buf ^ 1][(NA + t.wave) * S + d][t.lane] = sb[d];
        }
    }
    if (!has1 || n1.v != cur.v) i8_flush<S>(t, cur.c, cur.v, acc);
    __syncthreads();
    cur = n1;
    buf ^= 1;
    return has1;
}

// C(m x n) = alpha A B^T + beta C.  One workgroup of 2 WM waves per
// (32 WM) x 64 tile, waves WM x 2 of 32 x 32 (2 x 2 MFMA blocks).  A stage is
// one (k-step, variant): the S digit fragments of the tile's 2 WM A and 4 B
// row blocks, double-buffered in LDS with one barrier per stage, the next
// stage fetched while this one multiplies.  Stages run chunk by chunk, and
// within a chunk variant by variant, so only one variant's S int32
// accumulators are live; after each (chunk, variant) they are scaled and
// added into C (FP64).
template <int S, int WM>
__global__ void __launch_bounds__(128 * WM) zgemm_i8_kernel(
    int m, int n, int nks, int nch, const i32x4* __restrict__ PA,
    const unsigned long long* __restrict__ MA, const i32x4* __restrict__ PB,
    const unsigned long long* __restrict__ MB, double2* __restrict__ C, int64_t ldc,
    double2 alpha, double2 beta) {
    constexpr int NA = 2 * WM;
    __shared__ i32x4 lds[2][(NA + 4) * S][64];
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
    t.tm = tile / nt;
    t.tn = tile % nt;
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

#pragma unroll
    for (int d = 0; d < S; ++d) {
        lds[0][t.wave * S + d][t.lane] = ga[d * 64];
        if (t.wave < 4) lds[0][(NA + t.wave) * S + d][t.lane] = gb[d * 64];
    }
    __syncthreads();

    i32x4 acc[S][2][2];
#pragma unroll
    for (int o = 0; o < S; ++o)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[o][i][j] = (i32x4){0, 0, 0, 0};

    I8Stage cur = {0, 0, 0};
    int buf = 0;
    while (i8_stage<S, WM>(t, lds, ga, gb, cur, buf, acc)) {
    }
}

// ---------------------------------------------------------------- host side
void i8_layout(I8Operand& S, int rows, int k, int digits) {
    S.rows = rows;
    S.k = k;
    S.digits = digits;
    S.rows_pad = (rows + 127) / 128 * 128;   // whole 128-row GEMM tiles
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
                      double2 alpha, double2 beta, hipStream_t st) {
    // must match the kernel's tile count; planes cover whole 128-row tiles
    const int64_t tiles = (int64_t)((A.rows + 32 * WM - 1) / (32 * WM)) * ((B.rows + 63) / 64);
    hipLaunchKernelGGL((zgemm_i8_kernel<S, WM>), dim3((unsigned)tiles), dim3(128 * WM), 0, st,
                       A.rows, B.rows, A.nks, A.nch, (const i32x4*)A.planes, A.mx,
                       (const i32x4*)B.planes, B.mx, C, ldc, alpha, beta);
}

int zgemm_i8(const I8Operand& A, const I8Operand& B, double2* C, int64_t ldc, double2 alpha,
             double2 beta, hipStream_t st) {
    if (A.nks != B.nks || A.digits != B.digits) return -1;
    const int m = A.rows, n = B.rows;
    if (m <= 0 || n <= 0) return 0;
    if (A.nks == 0) return -1;
    static const int wm = [] {
        const char* e = getenv("RCC_I8_WM");
        return e ? atoi(e) : 4;
    }();
    if (A.digits == 4) {
        if (wm == 2) launch_i8<4, 2>(A, B, C, ldc, alpha, beta, st);
        else launch_i8<4, 4>(A, B, C, ldc, alpha, beta, st);
    } else if (A.digits == 5) {
        if (wm == 2) launch_i8<5, 2>(A, B, C, ldc, alpha, beta, st);
        else launch_i8<5, 4>(A, B, C, ldc, alpha, beta, st);
    } else {
        return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace rcc
