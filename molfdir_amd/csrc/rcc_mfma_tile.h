// MFMA tile building blocks shared by the complex FP64 GEMM kernels
// (rcc_zgemm.hip, rcc_triples.hip): tile constants, the per-thread
// global -> register operand loaders and the register -> LDS stores.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "rcc_internal.h"

namespace rcc {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int BM = 64;
constexpr int BN = 64;
constexpr int BK = 16;
constexpr int PAD = 16;          // LDS row padding (doubles), breaks bank aliasing
constexpr int LDA_S = BM + PAD;
constexpr int LDB_S = BN + PAD;

// LDS column swizzle.  An operand whose global layout is k-fastest (A with op
// T/C, B with op N) is written to its [k][row] LDS tile by lanes that differ
// in k, which would put 16 lanes on the same banks; XOR-ing the in-row index
// with the k-row (bits 0-3 only) spreads them.  Readers apply the same map;
// the 16-element MFMA blocks stay contiguous because bit 4 and above are
// untouched.
template <bool SW>
__device__ __forceinline__ int swz(int krow, int col) {
    return SW ? (col ^ (krow & 15)) : col;
}

// Per-thread global -> register loader for one operand tile.  Each thread
// owns 4 elements of the BM x BK (A) or BK x BN (B) tile; their addresses are
// formed once and advanced by a constant stride per K-step.  Rows / columns
// past the matrix edge are clamped to the last valid one: they only feed
// output elements that are never stored.  Only the last K-step of a split
// checks k (elements past kend are zero and are not read).
struct Loader {
    const double2* p[4];
    int kofs[4];          // k offset of element j inside a K-step
    int64_t step;         // pointer advance per K-step (elements)
};

template <int OP>
__device__ __forceinline__ double2 cj(double2 v) {
    if (OP == OP_C || OP == OP_R) v.y = -v.y;
    return v;
}

// N and R read the operand as stored (R conjugates on load), T and C transposed
template <int OP>
constexpr bool as_stored() { return OP == OP_N || OP == OP_R; }

// A tile (BM x BK of op(A)); A stored m x k (OP_N / OP_R) or k x m (OP_T / OP_C).
template <int OPA>
__device__ __forceinline__ Loader make_loader_a(const double2* __restrict__ A, int64_t lda, int m,
                                                int m0, int kbeg, int tid) {
    Loader L;
    if (as_stored<OPA>()) {
        const int row = tid & 63, kc = tid >> 6;
        const int64_t gm = min(m0 + row, m - 1);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            L.kofs[j] = kc + 4 * j;
            L.p[j] = A + gm + (int64_t)(kbeg + kc + 4 * j) * lda;
        }
        L.step = (int64_t)BK * lda;
    } else {
        const int kk = tid & 15, row = tid >> 4;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t gm = min(m0 + row + 16 * j, m - 1);
            L.kofs[j] = kk;
            L.p[j] = A + (kbeg + kk) + gm * lda;
        }
        L.step = BK;
    }
    return L;
}

// B tile (BK x BN of op(B)); B stored k x n (OP_N / OP_R) or n x k (OP_T / OP_C).
template <int OPB>
__device__ __forceinline__ Loader make_loader_b(const double2* __restrict__ B, int64_t ldb, int n,
                                                int n0, int kbeg, int tid) {
    Loader L;
    if (as_stored<OPB>()) {
        const int kk = tid & 15, col = tid >> 4;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t gn = min(n0 + col + 16 * j, n - 1);
            L.kofs[j] = kk;
            L.p[j] = B + (kbeg + kk) + gn * ldb;
        }
        L.step = BK;
    } else {
        const int col = tid & 63, kc = tid >> 6;
        const int64_t gn = min(n0 + col, n - 1);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            L.kofs[j] = kc + 4 * j;
            L.p[j] = B + gn + (int64_t)(kbeg + kc + 4 * j) * ldb;
        }
        L.step = (int64_t)BK * ldb;
    }
    return L;
}

// Loads the K-step starting at k0 (relative offset kr = k0 - kbeg is implied
// by how far the pointers were advanced); full = every k of the step < kend.
template <int OP>
__device__ __forceinline__ void load_step(const Loader& L, int k0, int kend, bool full,
                                          double2 (&r)[4]) {
    if (full) {
#pragma unroll
        for (int j = 0; j < 4; ++j) r[j] = cj<OP>(L.p[j][0]);
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            r[j] = (k0 + L.kofs[j] < kend) ? cj<OP>(L.p[j][0]) : make_double2(0.0, 0.0);
    }
}

__device__ __forceinline__ void advance(Loader& L) {
#pragma unroll
    for (int j = 0; j < 4; ++j) L.p[j] += L.step;
}

template <int OPA>
__device__ __forceinline__ void store_a(double* __restrict__ sre, double* __restrict__ sim,
                                        int tid, const double2 (&ra)[4]) {
    if (as_stored<OPA>()) {
        const int row = tid & 63;
        const int kc = tid >> 6;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            sre[(kc + 4 * j) * LDA_S + row] = ra[j].x;
            sim[(kc + 4 * j) * LDA_S + row] = ra[j].y;
        }
    } else {
        const int kk = tid & 15;
        const int row = tid >> 4;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            sre[kk * LDA_S + swz<true>(kk, row + 16 * j)] = ra[j].x;
            sim[kk * LDA_S + swz<true>(kk, row + 16 * j)] = ra[j].y;
        }
    }
}

template <int OPB>
__device__ __forceinline__ void store_b(double* __restrict__ sre, double* __restrict__ sim,
                                        int tid, const double2 (&rb)[4]) {
    if (as_stored<OPB>()) {
        const int kk = tid & 15;
        const int col = tid >> 4;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            sre[kk * LDB_S + swz<true>(kk, col + 16 * j)] = rb[j].x;
            sim[kk * LDB_S + swz<true>(kk, col + 16 * j)] = rb[j].y;
        }
    } else {
        const int col = tid & 63;
        const int kc = tid >> 6;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            sre[(kc + 4 * j) * LDB_S + col] = rb[j].x;
            sim[(kc + 4 * j) * LDB_S + col] = rb[j].y;
        }
    }
}

// workgroups are dealt to the 8 XCDs round-robin
constexpr int NXCD = 8;

}  // namespace rcc
