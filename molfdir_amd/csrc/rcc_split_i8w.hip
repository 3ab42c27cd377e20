// Split-INT8 complex GEMM, wide tile: 128 x 128 per workgroup of 8 waves,
// 64 x 32 per wave.  Same digit planes, exponents, stage order and flush arithmetic as
// zgemm_i8_kernel (rcc_split_i8.hip), so every element of C receives the same
// products in the same order and the result is bit-identical; only the tile
// shape differs.
//
// Why a wider tile: with a 32 x 32 block per wave, one stage (k-step of 64,
// one variant) reads 4 S digit fragments from LDS for S (S + 1) / 2 x 4 MFMAs,
// and together with the DMA writes the LDS traffic per stage exceeds the
// matrix-core time.  A 64 x 32 block per wave reads 6 S fragments for
// S (S + 1) / 2 x 8 MFMAs, and the 128 x 128 workgroup tile halves the DMA
// bytes per product: about a quarter less LDS traffic per product in all.
// Its 8 blocks of S int32 accumulators take 128 registers for S = 4.  Two LDS
// stages of 16 S KiB (128 KiB for S = 4).
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>
#include <algorithm>
#include "rcc_internal.h"

namespace rcc {

namespace {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void* lds_ptr;

constexpr int KSTEP = 64;                 // k per MFMA step / fragment
constexpr int KS_PER_CH = 8192 / KSTEP;   // k-steps per scale chunk

__device__ __forceinline__ int chunk_exp(unsigned long long bits) {
    const double mx = __longlong_as_double((long long)bits);
    int e = 0;
    if (mx > 0.0) (void)frexp(mx, &e);
    return e;
}

template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// stage st = (chunk, variant, k-step) in chunk-major, then variant-major order
struct Stage {
    int c, v, ks;
    bool last;
};

__device__ __forceinline__ Stage stage_of(int st, int nks) {
    const int full = nks / KS_PER_CH, rem = nks % KS_PER_CH;
    Stage s;
    int r, len;
    if (st < full * 3 * KS_PER_CH) {
        s.c = st / (3 * KS_PER_CH);
        r = st % (3 * KS_PER_CH);
        len = KS_PER_CH;
    } else {
        s.c = full;
        r = st - full * 3 * KS_PER_CH;
        len = rem;
    }
    s.v = r / len;
    s.ks = s.c * KS_PER_CH + r % len;
    s.last = r % len == len - 1;
    return s;
}

struct Tile {
    int m, n, nch, lane, row0, col0;
    const unsigned long long* MA;
    const unsigned long long* MB;
    double2* C;
    int64_t ldc;
    double2 alpha, beta;
};

// (chunk, variant) complete: the flush of zgemm_i8_kernel on a 64 x 32 block
template <int S>
__device__ __forceinline__ void flush(const Tile& t, int c, int v, i32x4 (&acc)[S][4][2]) {
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
    for (int i = 0; i < 4; ++i)
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
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[o][i][j] = (i32x4){0, 0, 0, 0};
}

// DMA of stage st into an LDS stage: wave w copies the S digit fragments of
// A row block w and B row block w of the tile (2 S loads of 16 bytes per lane)
template <int S>
__device__ __forceinline__ void copy_stage(i32x4 (*dst)[64], const i32x4* ga, const i32x4* gb,
                                           int wave, int st, int nks) {
    const Stage g = stage_of(st, nks);
    const int64_t off = ((int64_t)g.ks * 3 + g.v) * S * 64;
#pragma unroll
    for (int d = 0; d < S; ++d) {
        __builtin_amdgcn_global_load_lds((const void*)(ga + off + d * 64),
                                         (lds_ptr)&dst[wave * S + d][0], 16, 0, 0);
        __builtin_amdgcn_global_load_lds((const void*)(gb + off + d * 64),
                                         (lds_ptr)&dst[(8 + wave) * S + d][0], 16, 0, 0);
    }
}

// C(m x n) = alpha A B^T + beta C; 512 threads, waves 2 x 4 of 64 x 32.
// Tile order as zgemm_i8_kernel (XCD ranges, bands of gm tile rows).
template <int S>
__global__ void __launch_bounds__(512) zgemm_i8w_kernel(
    int m, int n, int nks, int nch, const i32x4* __restrict__ PA,
    const unsigned long long* __restrict__ MA, const i32x4* __restrict__ PB,
    const unsigned long long* __restrict__ MB, double2* __restrict__ C, int64_t ldc,
    double2 alpha, double2 beta, int gm) {
    __shared__ i32x4 lds[2][16 * S][64];
    const int mt = (m + 127) / 128, nt = (n + 127) / 128;
    const int T = mt * nt;
    const int b = blockIdx.x;
    const int xcd = b % 8, idx = b / 8;
    const int q = T / 8, rr = T % 8;
    const int tile = (xcd < rr ? xcd * (q + 1) : rr * (q + 1) + (xcd - rr) * q) + idx;
    int tm, tn;
    {
        const int band = gm * nt;
        const int first = (tile / band) * gm;
        const int rows = min(mt - first, gm);
        const int r = tile % band;
        tm = first + r % rows;
        tn = r / rows;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;       // wn 0..3
    Tile t;
    t.m = m;
    t.n = n;
    t.nch = nch;
    t.lane = lane;
    t.row0 = tm * 128 + wm * 64;
    t.col0 = tn * 128 + wn * 32;
    t.MA = MA;
    t.MB = MB;
    t.C = C;
    t.ldc = ldc;
    t.alpha = alpha;
    t.beta = beta;
    const int64_t rbs = (int64_t)nks * 3 * S * 64;      // i32x4 per row block
    const i32x4* ga = PA + (8 * (int64_t)tm + wave) * rbs + lane;
    const i32x4* gb = PB + (8 * (int64_t)tn + wave) * rbs + lane;

    i32x4 acc[S][4][2];
#pragma unroll
    for (int o = 0; o < S; ++o)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[o][i][j] = (i32x4){0, 0, 0, 0};

    const int nst = 3 * nks;
    copy_stage<S>(lds[0], ga, gb, wave, 0, nks);
    for (int st = 0; st < nst; ++st) {
        wait_vmcnt<0>();                   // this wave's copies of stage st landed
        __builtin_amdgcn_s_barrier();      // everyone's copies landed; stage st - 1 read
        if (st + 1 < nst) copy_stage<S>(lds[(st + 1) & 1], ga, gb, wave, st + 1, nks);
        const int cb = st & 1;
        // the B fragments stay in registers; A is streamed one row block at a
        // time (each fragment is still read from LDS once per stage)
        i32x4 bl[2][S];
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int d = 0; d < S; ++d) bl[j][d] = lds[cb][(8 + 2 * wn + j) * S + d][lane];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            i32x4 a[S];
#pragma unroll
            for (int d = 0; d < S; ++d) a[d] = lds[cb][(4 * wm + i) * S + d][lane];
#pragma unroll
            for (int da = 0; da < S; ++da)
#pragma unroll
                for (int db = 0; db + da < S; ++db)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[da + db][i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                            a[da], bl[j][db], acc[da + db][i][j], 0, 0, 0);
        }
        const Stage g = stage_of(st, nks);
        if (g.last) flush<S>(t, g.c, g.v, acc);
    }
}

}  // namespace

// 128 x 128 tiles; the digit planes cover whole 128-row tiles (rows_pad is a
// multiple of 384).  Four digits only (two LDS stages of 64 KiB).
int launch_i8_wide(int S, int gm, const I8Operand& A, const I8Operand& B, double2* C, int64_t ldc,
                   double2 alpha, double2 beta, hipStream_t st) {
    if (S != 4) return -1;
    const int64_t tiles = (int64_t)((A.rows + 127) / 128) * ((B.rows + 127) / 128);
    hipLaunchKernelGGL((zgemm_i8w_kernel<4>), dim3((unsigned)tiles), dim3(512), 0, st, A.rows,
                       B.rows, A.nks, A.nch, (const i32x4*)A.planes, A.mx,
                       (const i32x4*)B.planes, B.mx, C, ldc, alpha, beta, std::max(gm, 1));
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace rcc
