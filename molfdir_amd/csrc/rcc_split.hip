// Split-BF16 complex GEMM: FP64 operands carried as three BF16 slices, products
// on BF16 matrix cores with FP32 accumulation over K-chunks of 128, chunk sums
// carried in FP64.
//
// Representation.  An operand X (R rows x K, complex) is carried as three real
// variants, Re X, Im X and Re X + Im X (the Gauss three-product form).  For
// each variant, each row and each chunk of 128 k there is a power-of-two scale
// 2^-e, e the frexp exponent of the chunk's largest magnitude, so the scaled
// values x lie in (-1, 1).  Each x is written as three BF16 slices by
// successive rounding,
//     x1 = bf16(x), x2 = bf16(x - x1), x3 = bf16(x - x1 - x2),
// which carries 24 significant bits (|x - x1 - x2 - x3| <= 2^-25 of the chunk
// maximum).
//
// Product.  For each variant the real product is sum_{i+j<=2} x_i y_j (the six
// slice products whose weight is at least 2^-16 of the leading one) on
// v_mfma_f32_16x16x32_bf16: every slice product is exact in FP32 and the MFMA
// sums them in FP32 over one 128-k chunk.  At the end of each (chunk, variant)
// the FP32 partials are scaled by 2^(eA+eB) and added to FP64 accumulators held
// in registers, Re += P0 - P1 and Im += P2 - P0 - P1, so FP32 rounding is
// confined to 128-term sums.
//
// Layout in HBM.  Slices are stored fragment-major: fragment (row block of 16,
// k-step of 32, variant, slice) is 1 KiB, lane l holding the 8 BF16 values of
// row 16 rb + l % 16, k = 32 ks + 8 (l / 16) + 0..7, which is exactly the
// lane's MFMA operand (A and B alike).  Exponents are bytes at
// ((chunk 3 + variant) rows_pad + row), so the exponents of one (chunk,
// variant) over a 128-row tile are 128 contiguous bytes.
//
// GEMM kernel.  A workgroup of 8 waves computes a 128 x 128 tile, each wave a
// 64 x 32 block (4 x 2 MFMA blocks of 16 x 16).  A stage is one (k-step,
// variant): the three slice fragments of the tile's 8 A and 8 B row blocks,
// 48 KiB, plus (at the last k-step of a chunk) the 256 bytes of that (chunk,
// variant)'s exponents.  Stages are copied global -> LDS by DMA two ahead into
// three LDS buffers; each wave waits only for its own copies before the stage
// barrier.  Per stage a wave reads 18 fragments from LDS for 48 MFMAs.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <cmath>
#include <cstdlib>
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>
#include "rcc_internal.h"

namespace rcc {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void* lds_ptr;

constexpr int BF_TILE = 128;                    // rows / columns per GEMM tile
constexpr int BF_KC = 128;                      // k per scale chunk (FP32 accumulation)
constexpr int BF_KSTEP = 32;                    // k per MFMA step / fragment
constexpr int BF_KS_PER_CH = BF_KC / BF_KSTEP;  // 4
constexpr int BF_SR = 16;                       // rows per slicing workgroup (one row block)

__device__ __forceinline__ double variant(double2 v, int which) {
    return which == 0 ? v.x : (which == 1 ? v.y : v.x + v.y);
}

__device__ __forceinline__ unsigned short bf_bits(__bf16 b) {
    return __builtin_bit_cast(unsigned short, b);
}

// ---------------------------------------------------------------- slicing
// One 256-thread workgroup per (BF_SR rows, one k chunk): the tile is read
// once (coalesced in either source layout) into LDS, the row maxima of the
// three variants give the exponents, then the tile's (row block, k-step)
// pairs x 3 variants x 3 slices fragments are written, 64 lanes per 1 KiB
// fragment.  One row block per workgroup keeps the LDS tile at 33 KB, so
// four workgroups share a CU (32 rows: 66 KB, two; 0.77 -> 0.52 ms for a
// 9000 x 9000 operand).
//
// Packed sources sliced without a full-size intermediate.  SRC 2, the ring
// form of T2: element (r, k) with r = a + nv i, k = c + nv kk is
// s sgn T2[ac, i kk] - t1(a, kk) t1(c, i), the sign + when a > c and i > kk
// agree; zero on a = c or i = kk.  SRC 3 and 4, the F1 expansion of a packed
// P (nvvt x ncol): F1(a, d + nv col) = sgn P[ad, col], + for a > d, zero on
// a = d; SRC 3 slices F1 itself (rows a), SRC 4 its transpose (rows d + nv col).
struct RingSrc {
    const double2* T2;
    const double2* t1;
    double s;
    int nv, no;
    int64_t nvvt;
};

__device__ __forceinline__ double2 f1_elem(const RingSrc& q, int a, int64_t j) {
    const int d = (int)(j % q.nv);
    const int64_t col = j / q.nv;
    double2 v = make_double2(0.0, 0.0);
    if (a > d) v = q.T2[pair_of(a, d, q.nv) + q.nvvt * col];
    else if (a < d) {
        v = q.T2[pair_of(d, a, q.nv) + q.nvvt * col];
        v.x = -v.x;
        v.y = -v.y;
    }
    return v;
}

__device__ __forceinline__ double2 ring_elem(const RingSrc& q, int r, int k) {
    const int a = r % q.nv, i = r / q.nv, c = k % q.nv, kk = k / q.nv;
    double2 v = make_double2(0.0, 0.0);
    if (a != c && i != kk) {
        const int64_t ac = a > c ? pair_of(a, c, q.nv) : pair_of(c, a, q.nv);
        const int64_t ik = i > kk ? pair_of(i, kk, q.no) : pair_of(kk, i, q.no);
        v = q.T2[ac + q.nvvt * ik];
        const double sg = ((a > c) == (i > kk)) ? q.s : -q.s;
        v.x *= sg;
        v.y *= sg;
    }
    if (q.t1) {
        const double2 x = q.t1[a + q.nv * kk], y = q.t1[c + q.nv * i];
        v.x -= x.x * y.x - x.y * y.y;
        v.y -= x.x * y.y + x.y * y.x;
    }
    return v;
}

// SRC 0: X[r + k ld]; 1: X[k + r ld]; 2 .. 4: packed sources (RingSrc)
template <int SRC>
__global__ void __launch_bounds__(256) bf_split_kernel(const double2* __restrict__ X, int64_t ld,
                                                       int R, int K, int conj, int nks,
                                                       int rows_pad, signed char* __restrict__ ex,
                                                       i32x4* __restrict__ P, RingSrc ring) {
    constexpr int TPR = 256 / BF_SR;        // threads per row for the maxima
    constexpr int KPT = BF_KC / TPR;        // k per thread
    __shared__ double2 t[BF_SR][BF_KC + 1];
    __shared__ int se[BF_SR][3];
    const int c = blockIdx.x, r0 = blockIdx.y * BF_SR, k0 = c * BF_KC;
    for (int e = threadIdx.x; e < BF_SR * BF_KC; e += 256) {
        int r, k;
        if (SRC != 1 && SRC != 4) { r = e % BF_SR; k = e / BF_SR; }   // rows fastest
        else { k = e % BF_KC; r = e / BF_KC; }                         // k fastest
        const int gr = r0 + r, gk = k0 + k;
        double2 v = make_double2(0.0, 0.0);
        if (gr < R && gk < K) {
            if (SRC == 2) v = ring_elem(ring, gr, gk);
            else if (SRC == 3) v = f1_elem(ring, gr, gk);
            else if (SRC == 4) v = f1_elem(ring, gk, gr);
            else v = SRC == 0 ? X[gr + (int64_t)gk * ld] : X[gk + (int64_t)gr * ld];
        }
        if (conj) v.y = -v.y;
        t[r][k] = v;
    }
    __syncthreads();
    {
        // TPR lanes per row, KPT k each, then a butterfly over the TPR
        const int r = threadIdx.x / TPR, part = threadIdx.x % TPR;
        double m0 = 0.0, m1 = 0.0, m2 = 0.0;
        for (int k = part * KPT; k < part * KPT + KPT; ++k) {
            const double2 v = t[r][k];
            m0 = fmax(m0, fabs(v.x));
            m1 = fmax(m1, fabs(v.y));
            m2 = fmax(m2, fabs(v.x + v.y));
        }
#pragma unroll
        for (int off = TPR / 2; off > 0; off >>= 1) {
            m0 = fmax(m0, __shfl_xor(m0, off, TPR));
            m1 = fmax(m1, __shfl_xor(m1, off, TPR));
            m2 = fmax(m2, __shfl_xor(m2, off, TPR));
        }
        if (part < 3) {
            const double m = part == 0 ? m0 : (part == 1 ? m1 : m2);
            int e = 0;
            if (m > 0.0) (void)frexp(m, &e);
            e = std::max(-127, std::min(127, e));
            se[r][part] = e;
            ex[((int64_t)c * 3 + part) * rows_pad + r0 + r] = (signed char)e;
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int pass = 0; pass < BF_SR / 16; ++pass) {
        const int pair = wave + 4 * pass;   // (row block, k-step) of this tile
        const int rbl = pair / BF_KS_PER_CH, ksl = pair % BF_KS_PER_CH;
        const int r = rbl * 16 + (lane & 15);
        const int kk = ksl * BF_KSTEP + 8 * (lane >> 4);
        const int64_t rb = (r0 >> 4) + rbl, ks = (int64_t)c * BF_KS_PER_CH + ksl;
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            const int e = se[r][v];
            unsigned int w[3][4];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                unsigned short h[3][2];
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    double x = ldexp(variant(t[r][kk + 2 * b + q], v), -e);
                    const __bf16 s1 = (__bf16)(float)x;
                    x -= (double)(float)s1;
                    const __bf16 s2 = (__bf16)(float)x;
                    x -= (double)(float)s2;
                    const __bf16 s3 = (__bf16)(float)x;
                    h[0][q] = bf_bits(s1);
                    h[1][q] = bf_bits(s2);
                    h[2][q] = bf_bits(s3);
                }
#pragma unroll
                for (int s = 0; s < 3; ++s) w[s][b] = (unsigned)h[s][0] | ((unsigned)h[s][1] << 16);
            }
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                const int64_t f = ((rb * nks + ks) * 3 + v) * 3 + s;
                P[f * 64 + lane] = (i32x4){(int)w[s][0], (int)w[s][1], (int)w[s][2], (int)w[s][3]};
            }
        }
    }
}

// ---------------------------------------------------------------- GEMM
// s_waitcnt vmcnt(N) only (expcnt and lgkmcnt at their no-wait maxima).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// stage st = (chunk, variant, k-step), chunk-major, then variant-major
struct BfStage {
    int c, v, ks;
    bool last;   // last k-step of its (chunk, variant)
};

__device__ __forceinline__ BfStage bf_stage_of(int st) {
    BfStage s;
    s.c = st / (3 * BF_KS_PER_CH);
    const int r = st % (3 * BF_KS_PER_CH);
    s.v = r / BF_KS_PER_CH;
    s.ks = s.c * BF_KS_PER_CH + r % BF_KS_PER_CH;
    s.last = r % BF_KS_PER_CH == BF_KS_PER_CH - 1;
    return s;
}

// Tile of the GEMM kernel: 128 x 32 WN, waves 2 x WN of 64 x 32, NB LDS
// buffers.  (4, 3): 128 x 128, 8 waves, one workgroup per CU, copies two
// stages ahead.  (2, 2): 128 x 64, 4 waves, two workgroups per CU whose stage
// barriers are independent, copies one stage ahead.
template <int WN>
struct BfShape {
    static constexpr int TN = 32 * WN;                  // tile columns
    static constexpr int WAVES = 2 * WN;
    static constexpr int NBB = TN / 16;                 // B row blocks per tile
    static constexpr int FRAGS = (8 + NBB) * 3;         // fragments per stage
    static constexpr int STAGE = FRAGS * 64 + 16;       // i32x4 per LDS stage (+ 256 exponent bytes)
    static constexpr int APW = 8 / WAVES;               // A row blocks copied per wave
};

// DMA of stage st into an LDS stage: wave w copies the three slice fragments of
// A row blocks w, w + WAVES, .. (ga, ga + WAVES rbs, ..) and of B row block w
// (16 bytes per lane and fragment); at the last k-step of a (chunk, variant)
// wave 0 also copies that (chunk, variant)'s exponents, lanes 0-31 those of
// the tile's 128 A rows and lanes 32-63 those of its B rows (4 bytes each; with
// 64 B rows, lanes 48-63 repeat lanes 32-47 into bytes the kernel does not
// read).  A row block that lies wholly past the matrix edge is not copied
// (bit a of `on` for A row block a of the wave, bit APW for its B row block):
// only waves that are idle in that tile would read it.  3 loads per copied
// row block, one more for wave 0 at those stages.
template <int WN>
__device__ __forceinline__ void bf_copy(i32x4* dst, const i32x4* ga, const i32x4* gb,
                                        int64_t rbs, const signed char* ge, int64_t ge_stride,
                                        int wave, int st, unsigned on) {
    using S = BfShape<WN>;
    const BfStage g = bf_stage_of(st);
    const int64_t off = ((int64_t)g.ks * 3 + g.v) * 3 * 64;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
#pragma unroll
        for (int a = 0; a < S::APW; ++a)
            if (on & (1u << a))
                __builtin_amdgcn_global_load_lds(
                    (const void*)(ga + (int64_t)a * S::WAVES * rbs + off + s * 64),
                    (lds_ptr)(dst + ((wave + a * S::WAVES) * 3 + s) * 64), 16, 0, 0);
        if (on & (1u << S::APW))
            __builtin_amdgcn_global_load_lds((const void*)(gb + off + s * 64),
                                             (lds_ptr)(dst + ((8 + wave) * 3 + s) * 64), 16, 0,
                                             0);
    }
    if (wave == 0 && g.last)
        __builtin_amdgcn_global_load_lds((const void*)(ge + ((int64_t)g.c * 3 + g.v) * ge_stride),
                                         (lds_ptr)(dst + S::FRAGS * 64), 4, 0, 0);
}

// s_waitcnt vmcnt(n) for a wave-uniform n of 0 .. 10 (0 above that: waits
// for every load, which is always safe)
__device__ __forceinline__ void wait_vm(int n) {
    switch (n) {
        case 1: wait_vmcnt<1>(); break;
        case 2: wait_vmcnt<2>(); break;
        case 3: wait_vmcnt<3>(); break;
        case 4: wait_vmcnt<4>(); break;
        case 5: wait_vmcnt<5>(); break;
        case 6: wait_vmcnt<6>(); break;
        case 7: wait_vmcnt<7>(); break;
        case 8: wait_vmcnt<8>(); break;
        case 9: wait_vmcnt<9>(); break;
        case 10: wait_vmcnt<10>(); break;
        default: wait_vmcnt<0>(); break;
    }
}

// this wave's copies of stage st have landed: with NB = 3 the copies of stage
// st + 1 (if any), `loads` of them (one more for wave 0's exponents), may
// still be in flight; with NB = 2 none is issued yet
template <int WN, int NB>
__device__ __forceinline__ void bf_wait(int wave, int st, int st1, int loads) {
    if (NB == 3 && st + 1 < st1) {
        if (wave == 0 && bf_stage_of(st + 1).last) wait_vm(loads + 1);
        else wait_vm(loads);
    } else {
        wait_vmcnt<0>();
    }
}

// C(m x n) = alpha A B^T + beta C, A: m rows, B: n rows, both split over the
// same k; 128 x 32 WN tiles, NB LDS buffers.  Tile order: each XCD
// (workgroups dealt round-robin) takes a contiguous range of tile numbers;
// gm > 1 numbers the tiles in bands of gm tile rows, column by column inside
// a band, so the workgroups an XCD runs together share A and B panels through
// that XCD's L2.  K-split: blockIdx.y is one of kparts equal ranges of k
// chunks, whose raw sum goes to W (part p at W + p m n, ld m) for a separate
// reduction; kparts = 1 writes C.
template <int WN, int NB>
__global__ void __launch_bounds__(128 * WN, 2) zgemm_bf_kernel(
    int m, int n, int nks, int rows_a, int rows_b, const i32x4* __restrict__ PA,
    const signed char* __restrict__ EA, const i32x4* __restrict__ PB,
    const signed char* __restrict__ EB, double2* __restrict__ C, int64_t ldc, double2 alpha,
    double2 beta, int gm, int kparts, double2* __restrict__ W) {
    using S = BfShape<WN>;
    static_assert(NB == 2 || NB == 3, "two or three LDS buffers");
    __shared__ i32x4 lds[NB][S::STAGE];
    const int mt = (m + BF_TILE - 1) / BF_TILE, nt = (n + S::TN - 1) / S::TN;
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
    const int wm = wave & 1, wn = wave >> 1;   // wn 0 .. WN - 1
    // a wave whose block lies wholly past the matrix edge only copies and syncs
    const bool active = tm * BF_TILE + wm * 64 < m && tn * S::TN + wn * 32 < n;
    const int64_t rbs = (int64_t)nks * 9 * 64;  // i32x4 per row block
    const i32x4* ga = PA + (8 * (int64_t)tm + wave) * rbs + lane;
    const i32x4* gb = PB + ((int64_t)S::NBB * tn + wave) * rbs + lane;
    const int eb_lane = (lane - 32) % (S::TN / 4);
    const signed char* ge = lane < 32 ? EA + (int64_t)tm * BF_TILE + 4 * lane
                                      : EB + (int64_t)tn * S::TN + 4 * eb_lane;
    const int64_t ge_stride = lane < 32 ? rows_a : rows_b;
    const int nkc = nks / BF_KS_PER_CH;
    const int part = blockIdx.y;
    const int st0 = 3 * BF_KS_PER_CH * (int)((int64_t)part * nkc / kparts);
    const int st1 = 3 * BF_KS_PER_CH * (int)((int64_t)(part + 1) * nkc / kparts);

    f32x4 acc[4][2];
    double cre[4][2][4], cim[4][2][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int r = 0; r < 4; ++r) cre[i][j][r] = cim[i][j][r] = 0.0;
        }

    // the row blocks this wave copies that start inside the matrix
    unsigned on = 0;
#pragma unroll
    for (int a = 0; a < S::APW; ++a)
        if (tm * BF_TILE + 16 * (wave + a * S::WAVES) < m) on |= 1u << a;
    if (tn * S::TN + 16 * wave < n) on |= 1u << S::APW;
    const int loads = 3 * __builtin_popcount(on);
    if (st0 < st1) bf_copy<WN>(lds[st0 % NB], ga, gb, rbs, ge, ge_stride, wave, st0, on);
    if (NB == 3 && st0 + 1 < st1)
        bf_copy<WN>(lds[(st0 + 1) % NB], ga, gb, rbs, ge, ge_stride, wave, st0 + 1, on);
    for (int st = st0; st < st1; ++st) {
        bf_wait<WN, NB>(wave, st, st1, loads);
        __builtin_amdgcn_s_barrier();
        // every wave is past stage st - 1, so its buffer can be refilled
        const int ahead = st + NB - 1;
        if (ahead < st1)
            bf_copy<WN>(lds[ahead % NB], ga, gb, rbs, ge, ge_stride, wave, ahead, on);
        const i32x4* L = lds[st % NB];
        if (active) {
            bf16x8 bl[2][3];
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int s = 0; s < 3; ++s)
                    bl[j][s] =
                        __builtin_bit_cast(bf16x8, L[((8 + 2 * wn + j) * 3 + s) * 64 + lane]);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                bf16x8 a[3];
#pragma unroll
                for (int s = 0; s < 3; ++s)
                    a[s] = __builtin_bit_cast(bf16x8, L[((4 * wm + i) * 3 + s) * 64 + lane]);
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    // slice pairs (si, sj) with si + sj <= 2, largest weights first
                    f32x4 x = acc[i][j];
                    x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bl[j][0], x, 0, 0, 0);
                    x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bl[j][1], x, 0, 0, 0);
                    x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], bl[j][0], x, 0, 0, 0);
                    x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bl[j][2], x, 0, 0, 0);
                    x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], bl[j][1], x, 0, 0, 0);
                    x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], bl[j][0], x, 0, 0, 0);
                    acc[i][j] = x;
                }
            }
            const BfStage g = bf_stage_of(st);
            if (g.last) {
                // (chunk, variant) complete: scale the FP32 partials by the row and
                // column exponents (the LDS copy E of this stage) and add them into
                // the FP64 accumulators, Re += P0 - P1, Im += P2 - P0 - P1 (the
                // factors are exact, so each FMA rounds as the plain sum)
                const unsigned char* E = (const unsigned char*)(L + S::FRAGS * 64);
                const double sre = g.v == 0 ? 1.0 : (g.v == 1 ? -1.0 : 0.0);
                const double sim = g.v == 2 ? 1.0 : -1.0;
                int eb[2];
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    eb[j] = (signed char)E[128 + wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int ea4 = *(const int*)(E + wm * 64 + i * 16 + 4 * (lane >> 4));
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int ea = (signed char)((ea4 >> (8 * r)) & 0xff);
                            const double p = ldexp((double)acc[i][j][r], ea + eb[j]);
                            cre[i][j][r] = fma(sre, p, cre[i][j][r]);
                            cim[i][j][r] = fma(sim, p, cim[i][j][r]);
                        }
                        acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
                    }
                }
            }
        }
    }
    // D layout of v_mfma_f32_16x16x32_bf16: lane l, element r -> row 4 (l / 16) + r, col l % 16
    double2* Wp = W ? W + (int64_t)part * m * n : nullptr;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = tm * BF_TILE + wm * 64 + i * 16 + 4 * (lane >> 4) + r;
                const int col = tn * S::TN + wn * 32 + j * 16 + (lane & 15);
                if (row < m && col < n) {
                    const double xr = cre[i][j][r], xi = cim[i][j][r];
                    if (Wp) {
                        Wp[row + (int64_t)col * m] = make_double2(xr, xi);
                        continue;
                    }
                    double2* cp = C + row + (int64_t)col * ldc;
                    double2 o;
                    o.x = alpha.x * xr - alpha.y * xi;
                    o.y = alpha.x * xi + alpha.y * xr;
                    if (beta.x != 0.0 || beta.y != 0.0) {
                        const double2 cv = *cp;
                        o.x += beta.x * cv.x - beta.y * cv.y;
                        o.y += beta.x * cv.y + beta.y * cv.x;
                    }
                    *cp = o;
                }
            }
}

// wn: tile width 32 wn (4: 128 x 128 with three LDS buffers; 2: 128 x 64 with
// two); gm: tile-row band; kparts: K-split
struct BfChoice {
    int wn, gm, kparts;
};

template <int WN, int NB>
void launch_bf_t(const SplitOperand& A, const SplitOperand& B, double2* C, int64_t ldc,
                 double2 alpha, double2 beta, int gm, int kp, double2* W, hipStream_t st) {
    // must match the kernel's tile count; the planes cover whole 128-row tiles
    const int64_t tiles = (int64_t)((A.rows + BF_TILE - 1) / BF_TILE) *
                          ((B.rows + BfShape<WN>::TN - 1) / BfShape<WN>::TN);
    hipLaunchKernelGGL((zgemm_bf_kernel<WN, NB>), dim3((unsigned)tiles, (unsigned)kp),
                       dim3(128 * WN), 0, st, A.rows, B.rows, A.nks, A.rows_pad, B.rows_pad,
                       (const i32x4*)A.planes, A.ex, (const i32x4*)B.planes, B.ex, C, ldc, alpha,
                       beta, gm, kp, W);
}

void launch_bf(const SplitOperand& A, const SplitOperand& B, double2* C, int64_t ldc,
               double2 alpha, double2 beta, BfChoice ch, double2* ws, hipStream_t st) {
    const int kp = std::max(ch.kparts, 1);
    const int gm = std::max(ch.gm, 1);
    double2* W = kp > 1 ? ws : nullptr;
    if (ch.wn == 2) launch_bf_t<2, 2>(A, B, C, ldc, alpha, beta, gm, kp, W, st);
    else launch_bf_t<4, 3>(A, B, C, ldc, alpha, beta, gm, kp, W, st);
    if (kp > 1) launch_split_reduce(A.rows, B.rows, kp, ws, C, ldc, alpha, beta, st);
}

// Tile, tile-order and K-split autotuner: the first product of a shape (m, n,
// k steps) met outside graph capture times both tiles with tile-row bands gm
// of 1, 2, 4 and 8, and k split into equal ranges of chunks (2 .. 12 and the
// splits that fill 1 .. 4 rounds of one 128 x 128 tile per CU, for products of
// fewer than two rounds of tiles, else 2 and 3; partial sums in ws, reduced in
// a fixed order) with bands 1 and 4, on a scratch C of the product's size
// (the live C is never touched), and keeps the fastest for the process.
// RCC_BF_TILE (128 or 64, the tile width), RCC_BF_GM and RCC_BF_KPARTS fix the
// choice; with the autotuner off it is (128, 1, 1).  Tile and band do not
// change the result; a K-split sums the chunk ranges in another order.
std::mutex bf_tune_mu;
std::map<std::tuple<int, int, int>, BfChoice> bf_tuned;

BfChoice bf_choose(const SplitOperand& A, const SplitOperand& B, double2* C, int64_t ldc,
                   double2* ws, int64_t ws_elems, hipStream_t st) {
    const int nkc = A.nks / BF_KS_PER_CH;
    const int64_t mn = (int64_t)A.rows * B.rows;
    auto fits = [&](int kp) { return kp == 1 || (kp <= nkc && ws && kp * mn <= ws_elems); };
    const char* et = getenv("RCC_BF_TILE");
    const char* eg = getenv("RCC_BF_GM");
    const char* ek = getenv("RCC_BF_KPARTS");
    if (et || eg || ek) {
        BfChoice c{et && atoi(et) == 64 ? 2 : 4, eg ? std::max(1, std::min(32, atoi(eg))) : 1,
                   ek ? std::max(1, atoi(ek)) : 1};
        if (!fits(c.kparts)) c.kparts = 1;
        return c;
    }
    const BfChoice dflt{4, 1, 1};
    if (!zgemm_autotune_enabled()) return dflt;
    std::lock_guard<std::mutex> lk(bf_tune_mu);
    const auto key = std::make_tuple(A.rows, B.rows, A.nks);
    auto it = bf_tuned.find(key);
    if (it != bf_tuned.end()) return it->second;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
        return dflt;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return dflt;
    const int64_t tiles = (int64_t)((A.rows + BF_TILE - 1) / BF_TILE) *
                          ((B.rows + BF_TILE - 1) / BF_TILE);
    std::vector<BfChoice> cand;
    for (int wn : {4, 2})
        for (int g : {1, 2, 4, 8}) cand.push_back({wn, g, 1});
    std::vector<int> kps = tiles < 512 ? std::vector<int>{2, 3, 4, 6, 8, 12} : std::vector<int>{2, 3};
    if (tiles < 512) {
        const int64_t cus = std::max<int64_t>(resident_slots() / 2, 1);
        for (int64_t r = 1; r <= 4; ++r) {
            const int kp = (int)(r * cus / std::max<int64_t>(tiles, 1));
            if (kp > 12 && std::find(kps.begin(), kps.end(), kp) == kps.end()) kps.push_back(kp);
        }
    }
    // trial products on a scratch C (m x n, ld m)
    double2* ct = nullptr;
    if (hipMalloc((void**)&ct, (size_t)mn * sizeof(double2)) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return dflt;
    }
    (void)hipMemsetAsync(ct, 0, (size_t)mn * sizeof(double2), st);
    for (int kp : kps)
        if (fits(kp))
            for (int wn : {4, 2})
                for (int g : {1, 4}) cand.push_back({wn, g, kp});
    BfChoice best = dflt;
    float best_ms = 1e30f;
    for (const BfChoice& c : cand) {
        float ms_min = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
            (void)hipEventRecord(e0, st);
            launch_bf(A, B, ct, A.rows, make_double2(0, 0), make_double2(1, 0), c, ws, st);
            if (hipGetLastError() != hipSuccess) break;
            (void)hipEventRecord(e1, st);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0) ms_min = std::min(ms_min, ms);
        }
        if (ms_min < best_ms) {
            best_ms = ms_min;
            best = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(st);
    (void)hipFree(ct);
    bf_tuned[key] = best;
    return best;
}

}  // namespace

// ---------------------------------------------------------------- host side
void split_layout(SplitOperand& S, int rows, int k) {
    S.rows = rows;
    S.k = k;
    S.rows_pad = (rows + BF_TILE - 1) / BF_TILE * BF_TILE;
    S.k_pad = ((int64_t)k + BF_KC - 1) / BF_KC * BF_KC;
    S.nkc = (int)(S.k_pad / BF_KC);
    S.nks = (int)(S.k_pad / BF_KSTEP);
    S.plane_bytes = (int64_t)S.rows_pad * S.k_pad * 18;   // 3 variants x 3 slices x 2 bytes
    S.ex_bytes = (int64_t)S.nkc * 3 * S.rows_pad;
}

int split_operand(const SplitOperand& S, const double2* X, int64_t ld, bool k_contiguous,
                  bool conj, hipStream_t st) {
    if (S.rows <= 0 || S.k <= 0) return 0;
    dim3 g((unsigned)S.nkc, (unsigned)(S.rows_pad / BF_SR));
    const int cj = conj ? 1 : 0;
    const RingSrc none{nullptr, nullptr, 0.0, 1, 1, 0};
    if (k_contiguous)
        hipLaunchKernelGGL(bf_split_kernel<1>, g, dim3(256), 0, st, X, ld, S.rows, S.k, cj, S.nks,
                           S.rows_pad, S.ex, (i32x4*)S.planes, none);
    else
        hipLaunchKernelGGL(bf_split_kernel<0>, g, dim3(256), 0, st, X, ld, S.rows, S.k, cj, S.nks,
                           S.rows_pad, S.ex, (i32x4*)S.planes, none);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int split_f1_operand(const SplitOperand& S, const double2* P, int nv, int64_t nvvt, int64_t ncol,
                     bool transposed, hipStream_t st) {
    if (S.rows <= 0 || S.k <= 0) return 0;
    const int64_t wide = (int64_t)nv * ncol;
    if (transposed ? ((int64_t)S.rows != wide || S.k != nv) : (S.rows != nv || (int64_t)S.k != wide))
        return -1;
    dim3 g((unsigned)S.nkc, (unsigned)(S.rows_pad / BF_SR));
    const RingSrc q{P, nullptr, 0.0, nv, 1, nvvt};
    if (transposed)
        hipLaunchKernelGGL(bf_split_kernel<4>, g, dim3(256), 0, st, (const double2*)nullptr,
                           (int64_t)0, S.rows, S.k, 0, S.nks, S.rows_pad, S.ex, (i32x4*)S.planes, q);
    else
        hipLaunchKernelGGL(bf_split_kernel<3>, g, dim3(256), 0, st, (const double2*)nullptr,
                           (int64_t)0, S.rows, S.k, 0, S.nks, S.rows_pad, S.ex, (i32x4*)S.planes, q);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int split_ring_operand(const SplitOperand& S, const double2* T2, const double2* t1, double s,
                       int nv, int no, int64_t nvvt, hipStream_t st) {
    if (S.rows <= 0 || S.k <= 0) return 0;
    if ((int64_t)S.rows != (int64_t)nv * no || (int64_t)S.k != (int64_t)nv * no) return -1;
    dim3 g((unsigned)S.nkc, (unsigned)(S.rows_pad / BF_SR));
    const RingSrc q{T2, t1, s, nv, no, nvvt};
    hipLaunchKernelGGL(bf_split_kernel<2>, g, dim3(256), 0, st, (const double2*)nullptr,
                       (int64_t)0, S.rows, S.k, 0, S.nks, S.rows_pad, S.ex, (i32x4*)S.planes, q);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int zgemm_split(const SplitOperand& A, const SplitOperand& B, double2* C, int64_t ldc,
                double2 alpha, double2 beta, double2* ws, int64_t ws_elems, hipStream_t st) {
    if (A.k_pad != B.k_pad) return -1;
    const int m = A.rows, n = B.rows;
    if (m <= 0 || n <= 0) return 0;
    const BfChoice ch = bf_choose(A, B, C, ldc, ws, ws_elems, st);
    launch_bf(A, B, C, ldc, alpha, beta, ch, ws, st);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace rcc
