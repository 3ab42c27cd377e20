// Complex FP64 MFMA tile core shared by the GEMM engine (rcc_zgemm.hip) and
// the triples GEMM (rcc_triples.hip).
//
// A workgroup of 256 threads (2 x 2 wave64) owns a (32 TI) x (32 TJ) tile
// of C; each wave a (16 TI) x (16 TJ) block of v_mfma_f64_16x16x4_f64
// results.  K is walked in stages of ZBK = 8 through two LDS buffers.  Every
// stage is copied global -> LDS directly (global_load_lds_dwordx4: one
// complex element, 16 bytes, per lane and copy; no register staging), and
// the copies of stage s + 1 run while the MFMAs of stage s read the other
// buffer.  Each buffer is its own __shared__ array and the stage loop is
// unrolled by two, so the compiler can tell the reads of one buffer from the
// copies in flight into the other and does not wait for those copies before
// reading.  Per stage a wave waits for its own copies (vmcnt 0) and then
// for the workgroup (s_barrier): after that every copy of the stage has
// landed and every wave has finished reading the buffer the next copies
// overwrite.
//
// Operands are kept in LDS as interleaved complex (re, im), so one
// ds_read_b128 gives a lane both parts of its MFMA operand.  Two layouts,
// chosen by how the operand is stored in memory:
//   row-fastest (its m or n index fastest in memory):  p = k R + r
//   k-fastest:  p = ZBK r + (k ^ zswz(r))
// where p counts complex elements of the stage, r the row (m or n) inside
// the tile and k the index inside the stage.  The XOR swizzle puts the 16
// rows a 16-lane group reads at one k into 16 distinct 16-byte bank groups.
// The copy lane -> slot map is the identity on p (the DMA writes 64
// consecutive slots per instruction), so the swizzle is applied to the
// global address each lane reads, which stays within one 128-byte run.
//
// The complex product is formed from real MFMA products either classically
// (4 products: re = sum ArBr - sum AiBi, im = sum (ArBi + AiBr)) or by Gauss's three-
// multiplication form (P1 = ArBr, P2 = AiBi, P3 = (Ar+Ai)(Br+Bi); re = P1 -
// P2, im = P3 - P1 - P2; the "3M" product, 3/4 of the matrix-core work; its
// rounding error bound is normwise, Higham, SIAM J. Matrix Anal. Appl. 13
// (1992) 681).  Conjugated operands are negated in registers after the LDS
// read.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "rcc_internal.h"

namespace rcc {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void* lds_ptr;

// workgroups are dealt to the 8 XCDs round-robin
constexpr int NXCD = 8;
constexpr int ZBK = 8;          // k per stage
constexpr int ZNT = 256;        // threads per workgroup

template <int TI_, int TJ_>
struct ZTile {
    static constexpr int TI = TI_, TJ = TJ_;
    static constexpr int BM = 32 * TI, BN = 32 * TJ;
    static constexpr int SA = BM * ZBK, SB = BN * ZBK;   // complex elements per stage
    static constexpr int LA = SA / ZNT, LB = SB / ZNT;    // copies per thread per stage
    static_assert(LA * ZNT == SA && LB * ZNT == SB, "tile / thread mismatch");
};
using ZTileGemm = ZTile<3, 2>;    // 96 x 64: the GEMM engine
using ZTileT3 = ZTile<2, 3>;      // 64 x 96: the triples (m = v, n = v(v-1)/2)

__device__ __forceinline__ int zswz(int r) { return (r >> 1) & 7; }

template <bool ROWFAST, int R>
__device__ __forceinline__ int zpos(int r, int k) {
    return ROWFAST ? k * R + r : ZBK * r + (k ^ zswz(r));
}

// (row, k) held by LDS slot p
template <bool ROWFAST, int R>
__device__ __forceinline__ void zslot(int p, int& r, int& k) {
    if (ROWFAST) {
        k = p / R;
        r = p - k * R;
    } else {
        r = p / ZBK;
        k = (p % ZBK) ^ zswz(r);
    }
}

template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// one 16-byte global -> LDS copy per lane; dst is uniform over the wave
__device__ __forceinline__ void zcopy(const double2* src, double2* dst) {
    __builtin_amdgcn_global_load_lds((const void*)src, (lds_ptr)dst, 16, 0, 0);
}

template <class TL>
struct ZAcc {
    v4d c0[TL::TI][TL::TJ], c1[TL::TI][TL::TJ], c2[TL::TI][TL::TJ];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int i = 0; i < TL::TI; ++i)
#pragma unroll
            for (int j = 0; j < TL::TJ; ++j) {
                c0[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
                c1[i][j] = c0[i][j];
                c2[i][j] = c0[i][j];
            }
    }
    // complex result of element r of block (i, j).  3M: c0 = sum ArBr,
    // c1 = sum (Ar+Ai)(Br+Bi), c2 = sum AiBi; 4M: c0 = sum ArBr,
    // c1 = sum (ArBi + AiBr), c2 = sum AiBi
    template <int GAUSS>
    __device__ __forceinline__ double2 value(int i, int j, int r) const {
        if (GAUSS) return make_double2(c0[i][j][r] - c2[i][j][r], c1[i][j][r] - c0[i][j][r] - c2[i][j][r]);
        return make_double2(c0[i][j][r] - c2[i][j][r], c1[i][j][r]);
    }
};

// D layout of v_mfma_f64_16x16x4_f64 on gfx950 (pinned by the lane-layout
// probe test): lane l, element r of a 16 x 16 block -> row l / 16 + 4 r,
// column l % 16.
template <class TL>
__device__ __forceinline__ int zrow(int wm, int i, int lane, int r) {
    return wm * 16 * TL::TI + i * 16 + (lane >> 4) + 4 * r;
}
template <class TL>
__device__ __forceinline__ int zcol(int wn, int j, int lane) {
    return wn * 16 * TL::TJ + j * 16 + (lane & 15);
}

// Operands of one stage in registers: both 4-deep k steps of the wave's
// A rows and B columns (conjugated where asked).
template <class TL>
struct ZOps {
    double2 a[2][TL::TI], b[2][TL::TJ];
};

// Reads one stage from LDS: A tile in `la` (row-fastest if AR), B tile in `lb`
// (its n index fastest if BR); ca / cb conjugate the operand.
template <class TL, bool AR, bool BR>
__device__ __forceinline__ void zread(ZOps<TL>& o, const double2* __restrict__ la,
                                      const double2* __restrict__ lb, int wm, int wn, int lane,
                                      bool ca, bool cb) {
    const int lr = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int kr = 4 * h + lq;
#pragma unroll
        for (int i = 0; i < TL::TI; ++i) {
            o.a[h][i] = la[zpos<AR, TL::BM>(wm * 16 * TL::TI + i * 16 + lr, kr)];
            if (ca) o.a[h][i].y = -o.a[h][i].y;
        }
#pragma unroll
        for (int j = 0; j < TL::TJ; ++j) {
            o.b[h][j] = lb[zpos<BR, TL::BN>(wn * 16 * TL::TJ + j * 16 + lr, kr)];
            if (cb) o.b[h][j].y = -o.b[h][j].y;
        }
    }
}

// The MFMAs of one stage on operands in registers.
template <class TL, int GAUSS>
__device__ __forceinline__ void zmma(ZAcc<TL>& acc, const ZOps<TL>& o) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const double2* a = o.a[h];
        const double2* b = o.b[h];
        if (GAUSS) {
            double as[TL::TI], bs[TL::TJ];
#pragma unroll
            for (int i = 0; i < TL::TI; ++i) as[i] = a[i].x + a[i].y;
#pragma unroll
            for (int j = 0; j < TL::TJ; ++j) bs[j] = b[j].x + b[j].y;
#pragma unroll
            for (int i = 0; i < TL::TI; ++i)
#pragma unroll
                for (int j = 0; j < TL::TJ; ++j) {
                    acc.c0[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].x, acc.c0[i][j], 0, 0, 0);
                    acc.c1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[i], bs[j], acc.c1[i][j], 0, 0, 0);
                    acc.c2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].y, b[j].y, acc.c2[i][j], 0, 0, 0);
                }
        } else {
#pragma unroll
            for (int i = 0; i < TL::TI; ++i)
#pragma unroll
                for (int j = 0; j < TL::TJ; ++j) {
                    acc.c0[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].x, acc.c0[i][j], 0, 0, 0);
                    acc.c1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].y, acc.c1[i][j], 0, 0, 0);
                    acc.c2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].y, b[j].y, acc.c2[i][j], 0, 0, 0);
                    acc.c1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].y, b[j].x, acc.c1[i][j], 0, 0, 0);
                }
        }
    }
}

// The stage loop over nst full stages.  issue(s, buf) starts the copies of
// stage s into buf; read(s, buf, ops) reads stage s into registers.  Per
// stage: wait for the stage's copies (own, then the workgroup's), read it,
// start the next stage's copies into the other buffer, run the MFMAs (a
// scheduling barrier keeps them ahead of the next stage's wait).  The
// reads come before the copies in program order, so no LDS read ever waits
// for a copy in flight, whatever the compiler can prove about aliasing.
// lds0 / lds1 are two distinct __shared__ arrays of the calling kernel.
template <class TL, int GAUSS, class Issue, class Read>
__device__ __forceinline__ void zloop(ZAcc<TL>& acc, int nst, double2* lds0, double2* lds1,
                                      Issue&& issue, Read&& read) {
    if (nst > 0) issue(0, lds0);
    auto step = [&](int s, double2* cur, double2* nxt) {
        wait_vmcnt<0>();                   // this wave's copies of stage s landed
        __builtin_amdgcn_s_barrier();      // everyone's landed; stage s - 1 fully read
        ZOps<TL> o;
        read(s, cur, o);
        if (s + 1 < nst) issue(s + 1, nxt);
        zmma<TL, GAUSS>(acc, o);
        // keep this stage's MFMAs ahead of the next stage's wait: hoisted
        // above them, the wait would expose the latency of the copies just
        // issued
        __builtin_amdgcn_sched_barrier(0);
    };
    int s = 0;
    for (; s + 1 < nst; s += 2) {
        step(s, lds0, lds1);
        step(s + 1, lds1, lds0);
    }
    if (s < nst) step(s, lds0, lds1);
}

}  // namespace rcc
