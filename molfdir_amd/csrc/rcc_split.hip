// Split-BF16 complex GEMM: FP64 operands carried as three BF16 slices, products
// on BF16 matrix cores with FP32 accumulation over K-chunks of 128, chunk sums
// carried in FP64.
//
// Representation.  An operand X (R rows x K, complex) is cut into blocks of
// 64 rows x 128 k.  Each block is scaled by a power of two 2^-e so its largest
// |Re| or |Im| lies in [0.5, 1), and every scaled value x is written as three
// BF16 slices by successive rounding,
//     x1 = bf16(x), x2 = bf16(x - x1), x3 = bf16(x - x1 - x2),
// which carries 24 significant bits (|x - x1 - x2 - x3| <= 2^-25 |x|).  The
// slices of Re and Im are stored as six planes, K-major (row r, k contiguous)
// so one 16-byte load gives a lane its 8 consecutive k of an MFMA operand.
// Rows are padded to 64 and K to 128 with zeros.
//
// Product.  For each 16 x 16 output block, Re x Re, Im x Im, Re x Im and
// Im x Re are each sum_{i+j<=4} x_i y_j (the six slice products whose weight
// is at least 2^-16 of the leading one) on v_mfma_f32_16x16x32_bf16.  Every
// K-chunk the four FP32 partials are combined (RR - II, RI + IR), rescaled by
// 2^(eA+eB) and added to FP64 accumulators, so FP32 rounding is confined to
// 128-term sums.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <cmath>
#include <algorithm>
#include "rcc_internal.h"

namespace rcc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double v4d_s __attribute__((ext_vector_type(4)));

constexpr int SP_RB = 64;     // rows per scale block (= GEMM tile edge)
constexpr int SP_KC = 128;    // k per scale block (= FP32 accumulation chunk)
constexpr int SP_BK = 32;     // k per MFMA step

__device__ __forceinline__ __bf16 to_bf16(float f) { return (__bf16)f; }

// ---------------------------------------------------------------- block exponents
// One 256-thread block per (row-block, k-chunk): e = frexp exponent of the
// block's largest |Re| / |Im| (0 for an all-zero block).
template <int SRC>
__global__ void __launch_bounds__(256) split_exp_kernel(const double2* __restrict__ X, int64_t ld,
                                                        int R, int K, int conj, int nkc,
                                                        signed char* __restrict__ ex) {
    (void)conj;
    const int rb = blockIdx.y, kc = blockIdx.x;
    const int r0 = rb * SP_RB, k0 = kc * SP_KC;
    double mx = 0.0;
    for (int t = threadIdx.x; t < SP_RB * SP_KC; t += 256) {
        int r, k;
        if (SRC == 0) { r = t % SP_RB; k = t / SP_RB; }   // row-contiguous source
        else { k = t % SP_KC; r = t / SP_KC; }            // k-contiguous source
        r += r0;
        k += k0;
        if (r < R && k < K) {
            const double2 v = SRC == 0 ? X[r + (int64_t)k * ld] : X[k + (int64_t)r * ld];
            mx = fmax(mx, fmax(fabs(v.x), fabs(v.y)));
        }
    }
    // wave maxima by shuffles, then the four waves' in LDS
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
    __shared__ double sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double m = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
        int e = 0;
        if (m > 0.0) (void)frexp(m, &e);
        ex[(int64_t)rb * nkc + kc] = (signed char)std::max(-127, std::min(127, e));
    }
}

// ---------------------------------------------------------------- slicing
// Tiles of 32 rows x 64 k through LDS so both the FP64 read (either source
// layout) and the BF16 plane writes (k contiguous) are coalesced.
template <int SRC>
__global__ void __launch_bounds__(256) split_planes_kernel(const double2* __restrict__ X,
                                                           int64_t ld, int R, int K, int conj,
                                                           int nkc, const signed char* __restrict__ ex,
                                                           __bf16* __restrict__ P, int64_t ldk,
                                                           int64_t plane) {
    __shared__ double2 t[32][65];
    const int r0 = blockIdx.y * 32, k0 = blockIdx.x * 64;
    for (int e = threadIdx.x; e < 32 * 64; e += 256) {
        int r, k;
        if (SRC == 0) { r = e % 32; k = e / 32; }
        else { k = e % 64; r = e / 64; }
        const int gr = r0 + r, gk = k0 + k;
        double2 v = make_double2(0.0, 0.0);
        if (gr < R && gk < K) v = SRC == 0 ? X[gr + (int64_t)gk * ld] : X[gk + (int64_t)gr * ld];
        if (conj) v.y = -v.y;
        t[r][k] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * 64; e += 256) {
        const int k = e % 64, r = e / 64;
        const int gr = r0 + r, gk = k0 + k;
        const int sc = ex[(int64_t)(gr / SP_RB) * nkc + gk / SP_KC];
        const double2 v = t[r][k];
        double parts[2] = {ldexp(v.x, -sc), ldexp(v.y, -sc)};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            double x = parts[c];
            const __bf16 s1 = to_bf16((float)x);
            x -= (double)(float)s1;
            const __bf16 s2 = to_bf16((float)x);
            x -= (double)(float)s2;
            const __bf16 s3 = to_bf16((float)x);
            const int64_t o = (int64_t)gr * ldk + gk;
            P[(c * 3 + 0) * plane + o] = s1;
            P[(c * 3 + 1) * plane + o] = s2;
            P[(c * 3 + 2) * plane + o] = s3;
        }
    }
}

// ---------------------------------------------------------------- GEMM
// One 256-thread workgroup per 64 x 64 output tile, 4 waves of 32 x 32 each
// (2 x 2 blocks of 16 x 16).  A planes: Mp x Kp (row-major, k contiguous);
// B planes: Np x Kp.  C is FP64 complex, column-major (ldc), tile rows / cols
// beyond m / n are not stored.  Operands are read straight from L2 into
// registers (no LDS): each lane's fragment of an MFMA operand is 16
// contiguous bytes.
__global__ void __launch_bounds__(256) zgemm_split_kernel(
    int m, int n, int nkc, const __bf16* __restrict__ PA, int64_t ldka, int64_t planeA,
    const signed char* __restrict__ exA, const __bf16* __restrict__ PB, int64_t ldkb,
    int64_t planeB, const signed char* __restrict__ exB, double2* __restrict__ C, int64_t ldc,
    double2 alpha, double2 beta) {
    const int mt = (m + 63) / 64, nt = (n + 63) / 64;
    const int T = mt * nt;
    const int b = blockIdx.x;
    const int xcd = b % 8, idx = b / 8;
    const int q = T / 8, rr = T % 8;
    const int start = xcd < rr ? xcd * (q + 1) : rr * (q + 1) + (xcd - rr) * q;
    const int tile = start + idx;
    const int tm = tile / nt, tn = tile % nt;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int lr = lane & 15, lq = lane >> 4;

    // fragment base pointers: row (A) / column (B) of this lane, k offset 8*lq
    const __bf16* pa[2];
    const __bf16* pb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        pa[i] = PA + (int64_t)(tm * 64 + wm * 32 + i * 16 + lr) * ldka + 8 * lq;
        pb[i] = PB + (int64_t)(tn * 64 + wn * 32 + i * 16 + lr) * ldkb + 8 * lq;
    }
    v4d_s cre[2][2], cim[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            cre[i][j] = (v4d_s){0.0, 0.0, 0.0, 0.0};
            cim[i][j] = (v4d_s){0.0, 0.0, 0.0, 0.0};
        }
    for (int kc = 0; kc < nkc; ++kc) {
        f32x4 rr_[2][2], ii_[2][2], ri_[2][2], ir_[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                rr_[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
                ii_[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
                ri_[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
                ir_[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
            }
        for (int ks = 0; ks < SP_KC; ks += SP_BK) {
            const int64_t ko = (int64_t)kc * SP_KC + ks;
            bf16x8 a[2][6], bb[2][6];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int p = 0; p < 6; ++p) {
                    a[i][p] = *(const bf16x8*)(pa[i] + p * planeA + ko);
                    bb[i][p] = *(const bf16x8*)(pb[i] + p * planeB + ko);
                }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    // slice pairs (si, sj) with si + sj <= 2 (0-based), i.e. weight >= 2^-16
#define SP6(acc, ca, cb)                                                                       \
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][ca + 0], bb[j][cb + 0], acc, 0, 0, 0); \
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][ca + 0], bb[j][cb + 1], acc, 0, 0, 0); \
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][ca + 1], bb[j][cb + 0], acc, 0, 0, 0); \
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][ca + 0], bb[j][cb + 2], acc, 0, 0, 0); \
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][ca + 1], bb[j][cb + 1], acc, 0, 0, 0); \
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][ca + 2], bb[j][cb + 0], acc, 0, 0, 0);
                    SP6(rr_[i][j], 0, 0)
                    SP6(ii_[i][j], 3, 3)
                    SP6(ri_[i][j], 0, 3)
                    SP6(ir_[i][j], 3, 0)
#undef SP6
                }
        }
        // rescale the chunk and add in FP64
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int ea = exA[(int64_t)tm * nkc + kc];
                const int eb = exB[(int64_t)tn * nkc + kc];
                const double s = ldexp(1.0, ea + eb);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    cre[i][j][r] += s * ((double)rr_[i][j][r] - (double)ii_[i][j][r]);
                    cim[i][j][r] += s * ((double)ri_[i][j][r] + (double)ir_[i][j][r]);
                }
            }
    }
    // D layout of v_mfma_f32_16x16x32_bf16: lane l, element r -> row 4*(l/16)+r, col l%16
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = tm * 64 + wm * 32 + i * 16 + 4 * lq + r;
                const int gn = tn * 64 + wn * 32 + j * 16 + lr;
                if (gm < m && gn < n) {
                    const double xr = cre[i][j][r], xi = cim[i][j][r];
                    double2* cp = C + gm + (int64_t)gn * ldc;
                    double2 o;
                    o.x = alpha.x * xr - alpha.y * xi;
                    o.y = alpha.x * xi + alpha.y * xr;
                    if (beta.x != 0.0 || beta.y != 0.0) {
                        const double2 c = *cp;
                        o.x += beta.x * c.x - beta.y * c.y;
                        o.y += beta.x * c.y + beta.y * c.x;
                    }
                    *cp = o;
                }
            }
}

// ---------------------------------------------------------------- host side
int split_operand(const SplitOperand& S, const double2* X, int64_t ld, bool k_contiguous,
                  bool conj, hipStream_t st) {
    const int R = S.rows, K = S.k;
    if (R <= 0 || K <= 0) return 0;
    dim3 ge(S.nkc, S.rows_pad / SP_RB);
    dim3 gp((unsigned)(S.k_pad / 64), (unsigned)(S.rows_pad / 32));
    __bf16* planes = (__bf16*)S.planes;
    if (k_contiguous) {
        hipLaunchKernelGGL(split_exp_kernel<1>, ge, dim3(256), 0, st, X, ld, R, K, conj ? 1 : 0,
                           S.nkc, S.ex);
        hipLaunchKernelGGL(split_planes_kernel<1>, gp, dim3(256), 0, st, X, ld, R, K, conj ? 1 : 0,
                           S.nkc, S.ex, planes, S.k_pad, S.plane);
    } else {
        hipLaunchKernelGGL(split_exp_kernel<0>, ge, dim3(256), 0, st, X, ld, R, K, conj ? 1 : 0,
                           S.nkc, S.ex);
        hipLaunchKernelGGL(split_planes_kernel<0>, gp, dim3(256), 0, st, X, ld, R, K, conj ? 1 : 0,
                           S.nkc, S.ex, planes, S.k_pad, S.plane);
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

void split_layout(SplitOperand& S, int rows, int k) {
    S.rows = rows;
    S.k = k;
    S.rows_pad = (rows + SP_RB - 1) / SP_RB * SP_RB;
    S.k_pad = ((int64_t)k + SP_KC - 1) / SP_KC * SP_KC;
    S.nkc = (int)(S.k_pad / SP_KC);
    S.plane = (int64_t)S.rows_pad * S.k_pad;
}

int zgemm_split(const SplitOperand& A, const SplitOperand& B, double2* C, int64_t ldc,
                double2 alpha, double2 beta, hipStream_t st) {
    if (A.k_pad != B.k_pad) return -1;
    const int m = A.rows, n = B.rows;
    if (m <= 0 || n <= 0) return 0;
    const int64_t tiles = (int64_t)(A.rows_pad / 64) * (B.rows_pad / 64);
    hipLaunchKernelGGL(zgemm_split_kernel, dim3((unsigned)tiles), dim3(256), 0, st, m, n, A.nkc,
                       (const __bf16*)A.planes, A.k_pad, A.plane, A.ex, (const __bf16*)B.planes,
                       B.k_pad, B.plane, B.ex, C, ldc, alpha, beta);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace rcc
