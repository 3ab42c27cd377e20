// Non-iterative triples corrections of relccsd's T3CORR (cceqns.F:1336-1863)
// on CDNA4 FP64 matrix cores (gfx950).
//
// For every occupied triple i > j > k (k outer, j, i inner: the order of the
// reference's loops, cceqns.F:1426-1442) and every virtual triple a > b > c:
//
//   U'(a;bc) = sum_e  T(ae;ij) conj<ek||bc> + T(ae;jk) conj<ei||bc> - T(ae;ik) conj<ej||bc>
//            + sum_l  <al||ij> T(bc;kl) + <al||jk> T(bc;il) - <al||ik> T(bc;jl)
//   U(abc)   = U'(a;bc) - U'(b;ac) + U'(c;ab)
//   Y2, Y3   = the disconnected terms (V T1 + T F, and TAU3 T1, antisymmetrised)
//   ET1 = -sum |U|^2 / D,  ET2 = -sum Re(conj U Y2) / D,  ET3 = sum Re(conj U Y3)
//
// The six GEMMs of the reference per triple (cceqns.F:1471-1561) are one
// GEMM here, over a K dimension made of six segments: the A operand of a
// triple (v x K) is packed once per triple from T2 and VOOO (signs folded
// in, every segment zero-padded to a multiple of 16, so that no 8-deep
// stage of the GEMM core straddles two segments), and the B operand is read
// in place:
//   - particle segments: rows (occ * v + e) of the VOVV class itself, which
//     relccsd stores as ((E,K), C>D) with E fastest (the reference sorts and
//     conjugates it first, SRT23 + CONJUGA, cceqns.F:1410-1416; here the
//     conjugate is taken on load);
//   - hole segments: T(bc; occ, l) from T2 expanded once to all ordered
//     occupied pairs (nvvt x o^2, the role of SRT22, cceqns.F:1408).
// Each stage lies in one segment, so the B copies switch source per stage,
// not per element.  Tiles are ordered M-fastest inside each XCD's range: the
// v/64 tiles that share a B panel run together and read it through one L2.
//
// The energy kernel takes 8 x 8 x 8 blocks of (a, b, c) with block indices
// A >= B >= C; the pair rows of VVOO, T2 and TAU3 and the T1 / FVO values its
// disconnected terms need are staged in LDS, U' is read in place (each
// block consumes whole 128-byte runs of U' columns).  Partial sums are
// written per workgroup and reduced in a fixed order: the result does not
// depend on scheduling.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <string>
#include <vector>
#include "rcc_internal.h"
#include "rcc_zgemm_core.h"

namespace rcc {

namespace {

// packed pair index of (p, q), p > q, in an n-set: q outer, p inner
__device__ __forceinline__ int64_t pidx(int p, int q, int n) {
    return (int64_t)q * n - ((int64_t)q * (q + 1)) / 2 + (p - q - 1);
}

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 zscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

// T2 expanded over ordered occupied pairs: T2F[bc + (o * no + l) * nvvt] = T(bc; o l).
__global__ void t3_expand_t2_kernel(int no, int64_t nvvt, const double2* __restrict__ T2,
                                    double2* __restrict__ T2F) {
    const int64_t total = nvvt * no * no;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total;
         x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t bc = x % nvvt;
        const int64_t ol = x / nvvt;
        const int o = (int)(ol / no), l = (int)(ol % no);
        double2 v = make_double2(0.0, 0.0);
        if (o > l) v = T2[bc + pidx(o, l, no) * nvvt];
        else if (o < l) {
            v = T2[bc + pidx(l, o, no) * nvvt];
            v.x = -v.x;
            v.y = -v.y;
        }
        T2F[x] = v;
    }
}

// A operand of triple z (v x Kp, ld v), Kp = 3 nvp + 3 nop:
//   [s nvp, s nvp + v):          T(a,e; p_s), sign -1 for s = 2
//   [3 nvp + s nop, ... + o):     <al||p_s>,  sign -1 for s = 2
// with p_0 = ij, p_1 = jk, p_2 = ik; padding columns are zero.
__global__ void t3_pack_a_kernel(int nv, int no, int nvp, int nop, int64_t nvvt, int64_t nvo,
                                 const int4* __restrict__ tri, const double2* __restrict__ T2,
                                 const double2* __restrict__ VOOO, double2* __restrict__ Apk) {
    const int z = blockIdx.y;
    const int4 t = tri[z];
    const int64_t pr[3] = {pidx(t.x, t.y, no), pidx(t.y, t.z, no), pidx(t.x, t.z, no)};
    const int64_t Kp = 3 * (int64_t)nvp + 3 * (int64_t)nop;
    const int64_t total = (int64_t)nv * Kp;
    double2* out = Apk + (int64_t)z * total;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total;
         x += (int64_t)gridDim.x * blockDim.x) {
        const int a = (int)(x % nv);
        const int64_t col = x / nv;
        double2 v = make_double2(0.0, 0.0);
        double sg = 1.0;
        if (col < 3 * (int64_t)nvp) {
            const int s = (int)(col / nvp), e = (int)(col - (int64_t)s * nvp);
            sg = s == 2 ? -1.0 : 1.0;
            if (e < nv && e != a) {
                if (a > e) v = T2[pidx(a, e, nv) + pr[s] * nvvt];
                else {
                    v = T2[pidx(e, a, nv) + pr[s] * nvvt];
                    sg = -sg;
                }
            }
        } else {
            const int64_t c2 = col - 3 * (int64_t)nvp;
            const int s = (int)(c2 / nop), l = (int)(c2 - (int64_t)s * nop);
            sg = s == 2 ? -1.0 : 1.0;
            if (l < no) v = VOOO[(int64_t)l * nv + a + pr[s] * nvo];
        }
        out[x] = zscale(v, sg);
    }
}

// U'_z (v x nvvt, ld v) = A_z (v x Kp) . B_z (Kp x nvvt) on the GEMM core of
// rcc_zgemm_core.h: a 64 x 96 tile per workgroup (2 x 2 waves of 32 x 48),
// 8-deep stages copied global -> LDS.  A_z is m-fastest in memory.  The B
// operand changes source with the segment a stage lies in: particle stages
// read VOVV rows (occ v + e) in place, k-fastest in memory, conjugated after
// the LDS read; hole stages read T2F rows (occ o + l), n-fastest.  Rows past
// a segment's end (its zero padding in A_z) are clamped to its last row.
using TLT = ZTileT3;

template <int GAUSS>
__global__ void __launch_bounds__(256, 2)
t3_gemm_kernel(int nv, int no, int nvp, int nop, int nvvt, int64_t nvo,
               const int4* __restrict__ tri, const double2* __restrict__ Apk,
               const double2* __restrict__ VOVV, const double2* __restrict__ T2F,
               double2* __restrict__ U) {
    __shared__ double2 lds0[TLT::SA + TLT::SB];
    __shared__ double2 lds1[TLT::SA + TLT::SB];
    constexpr int BM = TLT::BM, BN = TLT::BN;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int m = nv, n = nvvt;
    const int Kp = 3 * nvp + 3 * nop;
    const int mt = (m + BM - 1) / BM, nt = (n + BN - 1) / BN;
    const int T = mt * nt;
    const int b = blockIdx.x;
    const int xcd = b % NXCD, idx = b / NXCD;
    const int q = T / NXCD, rr = T % NXCD;
    const int tile = (xcd < rr ? xcd * (q + 1) : rr * (q + 1) + (xcd - rr) * q) + idx;
    const int m0 = (tile % mt) * BM;      // M-fastest: tiles sharing a B panel are adjacent
    const int n0 = (tile / mt) * BN;
    const int z = blockIdx.y;
    const int4 t = tri[z];
    const int occ[3] = {t.z, t.x, t.y};   // odd occupied index of the pairs ij, jk, ik

    const double2* Az = Apk + (int64_t)z * m * Kp;
    int64_t offa[TLT::LA];
#pragma unroll
    for (int j = 0; j < TLT::LA; ++j) {
        int r, kq;
        zslot<true, BM>((j * 4 + wave) * 64 + lane, r, kq);
        offa[j] = min(m0 + r, m - 1) + (int64_t)kq * m;
    }
    // B slots in both layouts: (k, column offset) of each copy
    int kp[TLT::LB], kh[TLT::LB];
    int64_t cp[TLT::LB], ch[TLT::LB];
#pragma unroll
    for (int j = 0; j < TLT::LB; ++j) {
        const int p = (j * 4 + wave) * 64 + lane;
        int r;
        zslot<false, BN>(p, r, kp[j]);
        cp[j] = (int64_t)min(n0 + r, n - 1) * nvo;
        zslot<true, BN>(p, r, kh[j]);
        ch[j] = min(n0 + r, n - 1);
    }

    ZAcc<TLT> acc;
    acc.zero();
    zloop<TLT, GAUSS>(
        acc, Kp / ZBK, lds0, lds1,
        [&](int s, double2* buf) {
            const int k0 = s * ZBK;
#pragma unroll
            for (int j = 0; j < TLT::LA; ++j)
                zcopy(Az + (int64_t)k0 * m + offa[j], buf + (j * 4 + wave) * 64);
            double2* lb = buf + TLT::SA;
            if (k0 < 3 * nvp) {
                const int sg = k0 / nvp, e0 = k0 - sg * nvp;
                const double2* base = VOVV + (int64_t)occ[sg] * nv;
#pragma unroll
                for (int j = 0; j < TLT::LB; ++j)
                    zcopy(base + min(e0 + kp[j], nv - 1) + cp[j], lb + (j * 4 + wave) * 64);
            } else {
                const int c2 = k0 - 3 * nvp;
                const int sg = c2 / nop, l0 = c2 - sg * nop;
                const double2* base = T2F + (int64_t)occ[sg] * no * n;
#pragma unroll
                for (int j = 0; j < TLT::LB; ++j)
                    zcopy(base + (int64_t)min(l0 + kh[j], no - 1) * n + ch[j],
                          lb + (j * 4 + wave) * 64);
            }
        },
        [&](int s, double2* buf, ZOps<TLT>& o) {
            if (s * ZBK < 3 * nvp)
                zread<TLT, true, false>(o, buf, buf + TLT::SA, wm, wn, lane, false, true);
            else
                zread<TLT, true, true>(o, buf, buf + TLT::SA, wm, wn, lane, false, false);
        });

    double2* Uz = U + (int64_t)z * m * n;
#pragma unroll
    for (int i = 0; i < TLT::TI; ++i)
#pragma unroll
        for (int j = 0; j < TLT::TJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + zrow<TLT>(wm, i, lane, r);
                const int gn = n0 + zcol<TLT>(wn, j, lane);
                if (gm < m && gn < n) Uz[gm + (int64_t)gn * m] = acc.template value<GAUSS>(i, j, r);
            }
}

constexpr int EB = 8;   // energy block edge (a, b, c)

// One 8 x 8 x 8 block (A >= B >= C) of virtual triples of occupied triple z.
// Writes the block's three partial sums to part[(z * nblk + block) * 3 ..].
__global__ void __launch_bounds__(256)
t3_energy_kernel(int nv, int no, int64_t nvvt, int64_t nvo, const int4* __restrict__ tri,
                 const int4* __restrict__ blk, int64_t nblk, const double2* __restrict__ U,
                 const double2* __restrict__ T1, const double2* __restrict__ FVO,
                 const double2* __restrict__ T2, const double2* __restrict__ VVOO,
                 const double* __restrict__ eps, double* __restrict__ part) {
    // G[pt][pl][q]: pair type pt (0: (a,b), 1: (b,c), 2: (a,c)), local pair
    // pl = x_l + 8 y_l, q = VVOO(p_s) (0..2), T2(p_s) (3..5), TAU3(p_s) (6..8)
    __shared__ double2 G[3][EB * EB][9];
    // H[t][x_l][q]: index block t (0 a, 1 b, 2 c), q = T1(x, o_s), FVO(x, o_s)
    // with o = (k, i, j) and the sign of s = 2 folded in
    __shared__ double2 H[3][EB][6];
    __shared__ double red[4][3];

    const int tid = threadIdx.x;
    const int z = blockIdx.y;
    const int4 bb = blk[blockIdx.x];
    const int4 t = tri[z];
    const int i = t.x, j = t.y, k = t.z;
    const int64_t pr[3] = {pidx(i, j, no), pidx(j, k, no), pidx(i, k, no)};
    const int po[3][2] = {{i, j}, {j, k}, {i, k}};
    const int oc[3] = {k, i, j};
    const int base[3] = {bb.x * EB, bb.y * EB, bb.z * EB};

    if (tid < 3 * EB * 6) {
        const int tt = tid / (EB * 6), rem = tid % (EB * 6);
        const int xl = rem / 6, qq = rem % 6;
        const int x = base[tt] + xl;
        double2 v = make_double2(0.0, 0.0);
        if (x < nv) {
            const int s = qq % 3;
            const double2* src = qq < 3 ? T1 : FVO;
            v = src[(int64_t)oc[s] * nv + x];
            if (s == 2) v = zscale(v, -1.0);
        }
        H[tt][xl][qq] = v;
    }
    for (int e = tid; e < 3 * EB * EB * 9; e += 256) {
        const int pt = e / (EB * EB * 9), rem = e % (EB * EB * 9);
        const int pl = rem / 9, qq = rem % 9;
        const int xb = pt == 1 ? base[1] : base[0];
        const int yb = pt == 0 ? base[1] : base[2];
        const int x = xb + (pl & 7), y = yb + (pl >> 3);
        double2 v = make_double2(0.0, 0.0);
        if (x < nv && x > y) {
            const int s = qq % 3;
            const int64_t p = pidx(x, y, nv) + pr[s] * nvvt;
            if (qq < 3) v = VVOO[p];
            else {
                v = T2[p];
                if (qq >= 6) {
                    // TAU3 = T2 + (T1(x,o1) T1(y,o2) - T1(x,o2) T1(y,o1)) / 3
                    // (GETTAU3, ccintm.F:1055-1153)
                    const int o1 = po[s][0], o2 = po[s][1];
                    const double2 d1 = zmul(zscale(T1[(int64_t)o2 * nv + y], 1.0 / 3.0),
                                            T1[(int64_t)o1 * nv + x]);
                    const double2 d2 = zmul(zscale(T1[(int64_t)o1 * nv + y], 1.0 / 3.0),
                                            T1[(int64_t)o2 * nv + x]);
                    v.x += d1.x - d2.x;
                    v.y += d1.y - d2.y;
                }
            }
        }
        G[pt][pl][qq] = v;
    }
    __syncthreads();

    const double eo = eps[i] + eps[j] + eps[k];
    const double2* Uz = U + (int64_t)z * nv * nvvt;
    double e1 = 0.0, e2 = 0.0, e3 = 0.0;
    const int al = tid & 7, bl = (tid >> 3) & 7;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int cl = (tid >> 6) + 4 * r;
        const int a = base[0] + al, b = base[1] + bl, c = base[2] + cl;
        if (a < nv && a > b && b > c) {
            const double2 u1 = Uz[a + pidx(b, c, nv) * nv];
            const double2 u2 = Uz[b + pidx(a, c, nv) * nv];
            const double2 u3 = Uz[c + pidx(a, b, nv) * nv];
            const double2 u = make_double2(u1.x - u2.x + u3.x, u1.y - u2.y + u3.y);
            // Y'(xy; w) = sum_s G(xy; s) H(w; s) for the three (pair, index) slots
            double2 y2 = make_double2(0.0, 0.0), y3 = make_double2(0.0, 0.0);
            const int slot_pl[3] = {al + 8 * bl, bl + 8 * cl, al + 8 * cl};
            const int slot_w[3] = {2, 0, 1};
            const int slot_x[3] = {cl, al, bl};
            const double slot_sg[3] = {1.0, 1.0, -1.0};
#pragma unroll
            for (int sl = 0; sl < 3; ++sl) {
                const double2* g = G[sl][slot_pl[sl]];
                const double2* hh = H[slot_w[sl]][slot_x[sl]];
                double2 a2 = make_double2(0.0, 0.0), a3 = make_double2(0.0, 0.0);
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    const double2 p1 = zmul(g[s], hh[s]);
                    const double2 p2 = zmul(g[3 + s], hh[3 + s]);
                    const double2 p3 = zmul(g[6 + s], hh[s]);
                    a2.x += p1.x + p2.x;
                    a2.y += p1.y + p2.y;
                    a3.x += p3.x;
                    a3.y += p3.y;
                }
                y2.x += slot_sg[sl] * a2.x;
                y2.y += slot_sg[sl] * a2.y;
                y3.x += slot_sg[sl] * a3.x;
                y3.y += slot_sg[sl] * a3.y;
            }
            const double d = eps[no + a] + eps[no + b] + eps[no + c] - eo;
            e1 -= (u.x * u.x + u.y * u.y) / d;
            e2 -= (u.x * y2.x + u.y * y2.y) / d;
            e3 += u.x * y3.x + u.y * y3.y;
        }
    }
    // workgroup sums in a fixed order
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        e1 += __shfl_down(e1, off, 64);
        e2 += __shfl_down(e2, off, 64);
        e3 += __shfl_down(e3, off, 64);
    }
    if ((tid & 63) == 0) {
        red[tid >> 6][0] = e1;
        red[tid >> 6][1] = e2;
        red[tid >> 6][2] = e3;
    }
    __syncthreads();
    if (tid < 3) {
        const double s = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
        part[((int64_t)z * nblk + blockIdx.x) * 3 + tid] = s;
    }
}

// Sums n partial triples (x, y, z) of part into out[0..2], fixed order.
__global__ void __launch_bounds__(1024)
t3_reduce_kernel(const double* __restrict__ part, int64_t n, double* __restrict__ out) {
    __shared__ double sh[16][3];
    const int tid = threadIdx.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int64_t x = tid; x < n; x += 1024) {
        s0 += part[x * 3];
        s1 += part[x * 3 + 1];
        s2 += part[x * 3 + 2];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_down(s0, off, 64);
        s1 += __shfl_down(s1, off, 64);
        s2 += __shfl_down(s2, off, 64);
    }
    if ((tid & 63) == 0) {
        sh[tid >> 6][0] = s0;
        sh[tid >> 6][1] = s1;
        sh[tid >> 6][2] = s2;
    }
    __syncthreads();
    if (tid < 3) {
        double s = 0.0;
        for (int w = 0; w < 16; ++w) s += sh[w][tid];
        out[tid] = s;
    }
}

int round16(int x) { return (x + 15) / 16 * 16; }

}  // namespace

struct T3State {
    bool ready = false;            // every buffer below allocated for (no, nv, budget)
    int no = 0, nv = 0, zmax = 0;
    int64_t ntri = 0, nblk = 0, budget = 0;
    int4* tri = nullptr;
    int4* blk = nullptr;
    double2 *T2F = nullptr, *Apk = nullptr, *U = nullptr;
    double *part = nullptr, *sums = nullptr;
    int64_t nsums = 0;
    std::vector<hipEvent_t> ev;     // per chunk: GEMM begin / end, energy end
};

static void t3_release(T3State* s) {
    if (!s) return;
    for (void* p : {(void*)s->tri, (void*)s->blk, (void*)s->T2F, (void*)s->Apk, (void*)s->U,
                    (void*)s->part, (void*)s->sums})
        if (p) (void)hipFree(p);
    for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
    s->ev.clear();
    s->tri = nullptr;
    s->blk = nullptr;
    s->T2F = s->Apk = s->U = nullptr;
    s->part = s->sums = nullptr;
    s->ready = false;
}

void t3_free(T3State* s) {
    t3_release(s);
    delete s;
}

int64_t t3_default_budget() {
    const char* e = getenv("RCC_T3_BUDGET_MB");
    const int64_t mb = e ? atoll(e) : 4096;
    return std::max<int64_t>(mb, 1) << 20;
}

// On failure every buffer is released, so that a later call re-allocates
// instead of finding a half-built state, and the runtime's last-error state
// is cleared, so that the failure is not reported again by a later check.
#define T3CK(call)                                                              \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) {                                                 \
            err = std::string(#call) + ": " + hipGetErrorString(e_);            \
            (void)hipGetLastError();                                            \
            t3_release(s);                                                      \
            return -1;                                                          \
        }                                                                       \
    } while (0)

static int t3_prepare(T3State* s, int no, int nv, std::string& err) {
    if (s->ready && s->no == no && s->nv == nv && s->budget == t3_default_budget()) return 0;
    t3_release(s);
    s->no = no;
    s->nv = nv;
    s->budget = t3_default_budget();
    std::vector<int4> tri;
    for (int k = 0; k < no; ++k)
        for (int j = k + 1; j < no; ++j)
            for (int i = j + 1; i < no; ++i) tri.push_back(make_int4(i, j, k, 0));
    s->ntri = (int64_t)tri.size();
    const int nb = (nv + EB - 1) / EB;
    std::vector<int4> blk;
    for (int C = 0; C < nb; ++C)
        for (int B = C; B < nb; ++B)
            for (int A = B; A < nb; ++A) blk.push_back(make_int4(A, B, C, 0));
    s->nblk = (int64_t)blk.size();
    const int64_t nvvt = (int64_t)nv * (nv - 1) / 2;
    const int64_t per = (int64_t)nv * nvvt * 16 +
                        (int64_t)nv * (3 * round16(nv) + 3 * round16(no)) * 16;
    // grid.y holds the triples of a launch
    s->zmax = (int)std::max<int64_t>(
        1, std::min<int64_t>({s->ntri, s->budget / std::max<int64_t>(per, 1), 16384}));
    if (s->ntri == 0 || s->nblk == 0) {
        s->ready = true;
        return 0;
    }
    T3CK(hipMalloc((void**)&s->tri, tri.size() * sizeof(int4)));
    T3CK(hipMemcpy(s->tri, tri.data(), tri.size() * sizeof(int4), hipMemcpyHostToDevice));
    T3CK(hipMalloc((void**)&s->blk, blk.size() * sizeof(int4)));
    T3CK(hipMemcpy(s->blk, blk.data(), blk.size() * sizeof(int4), hipMemcpyHostToDevice));
    T3CK(hipMalloc((void**)&s->T2F, (size_t)nvvt * no * no * sizeof(double2)));
    const int64_t Kp = 3 * (int64_t)round16(nv) + 3 * (int64_t)round16(no);
    T3CK(hipMalloc((void**)&s->Apk, (size_t)s->zmax * nv * Kp * sizeof(double2)));
    T3CK(hipMalloc((void**)&s->U, (size_t)s->zmax * nv * nvvt * sizeof(double2)));
    T3CK(hipMalloc((void**)&s->part, (size_t)s->zmax * s->nblk * 3 * sizeof(double)));
    s->nsums = (s->ntri + s->zmax - 1) / s->zmax;
    T3CK(hipMalloc((void**)&s->sums, (size_t)s->nsums * 3 * sizeof(double)));
    s->ready = true;
    return 0;
}

T3State* t3_new() { return new T3State(); }

int t3_chunk(const T3State* s) { return s ? s->zmax : 0; }
int64_t t3_triples(int no) { return (int64_t)no * (no - 1) * (no - 2) / 6; }

int t3_run(T3State* s, const T3Input& in, int64_t first, int64_t count, hipStream_t st,
           double* et, T3Timing* tm, std::string& err) {
    et[0] = et[1] = et[2] = 0.0;
    if (tm) *tm = T3Timing{};
    const int no = in.no, nv = in.nv;
    if (no < 3 || nv < 3) return 0;
    if (t3_prepare(s, no, nv, err)) return -1;
    if (first < 0 || count < 0 || first + count > s->ntri) {
        err = "rcc_t3corr: triple range [" + std::to_string(first) + ", " +
              std::to_string(first + count) + ") outside [0, " + std::to_string(s->ntri) + ")";
        return -1;
    }
    if (count == 0) return 0;
    const int64_t nvvt = (int64_t)nv * (nv - 1) / 2, nvo = (int64_t)nv * no;
    if (nvvt > INT32_MAX / 2) {
        err = "rcc_t3corr: virtual pair count exceeds the kernel's 32-bit column index";
        return -1;
    }
    const int nvp = round16(nv), nop = round16(no);
    const int64_t Kp = 3 * (int64_t)nvp + 3 * (int64_t)nop;
    const int64_t nchunk = (count + s->zmax - 1) / s->zmax;
    while ((int64_t)s->ev.size() < 3 * nchunk + 2) {
        hipEvent_t e;
        T3CK(hipEventCreate(&e));
        s->ev.push_back(e);
    }
    hipEvent_t e_beg = s->ev[3 * nchunk], e_end = s->ev[3 * nchunk + 1];
    T3CK(hipEventRecord(e_beg, st));
    {
        const int64_t tot = nvvt * no * no;
        const int grid = (int)std::min<int64_t>((tot + 255) / 256, 65536);
        t3_expand_t2_kernel<<<grid, 256, 0, st>>>(no, nvvt, in.T2, s->T2F);
        T3CK(hipGetLastError());
    }
    const int mt = (nv + TLT::BM - 1) / TLT::BM, nt = (int)((nvvt + TLT::BN - 1) / TLT::BN);
    for (int64_t c = 0; c < nchunk; ++c) {
        const int64_t z0 = first + c * s->zmax;
        const int zn = (int)std::min<int64_t>(s->zmax, first + count - z0);
        const int4* tri = s->tri + z0;
        {
            const int64_t tot = (int64_t)nv * Kp;
            dim3 grid((unsigned)std::min<int64_t>((tot + 255) / 256, 4096), zn);
            t3_pack_a_kernel<<<grid, 256, 0, st>>>(nv, no, nvp, nop, nvvt, nvo, tri, in.T2, in.VOOO,
                                                   s->Apk);
            T3CK(hipGetLastError());
        }
        T3CK(hipEventRecord(s->ev[3 * c], st));
        dim3 gg((unsigned)(mt * nt), zn);
        if (in.gauss)
            t3_gemm_kernel<1><<<gg, 256, 0, st>>>(nv, no, nvp, nop, (int)nvvt, nvo, tri, s->Apk,
                                                  in.VOVV, s->T2F, s->U);
        else
            t3_gemm_kernel<0><<<gg, 256, 0, st>>>(nv, no, nvp, nop, (int)nvvt, nvo, tri, s->Apk,
                                                  in.VOVV, s->T2F, s->U);
        T3CK(hipGetLastError());
        T3CK(hipEventRecord(s->ev[3 * c + 1], st));
        dim3 ge((unsigned)s->nblk, zn);
        t3_energy_kernel<<<ge, 256, 0, st>>>(nv, no, nvvt, nvo, tri, s->blk, s->nblk, s->U, in.T1,
                                             in.FVO, in.T2, in.VVOO, in.eps, s->part);
        T3CK(hipGetLastError());
        t3_reduce_kernel<<<1, 1024, 0, st>>>(s->part, (int64_t)zn * s->nblk, s->sums + 3 * c);
        T3CK(hipGetLastError());
        T3CK(hipEventRecord(s->ev[3 * c + 2], st));
    }
    T3CK(hipEventRecord(e_end, st));
    std::vector<double> h(3 * nchunk);
    T3CK(hipMemcpyAsync(h.data(), s->sums, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    T3CK(hipStreamSynchronize(st));
    for (int64_t c = 0; c < nchunk; ++c)
        for (int q = 0; q < 3; ++q) et[q] += h[3 * c + q];
    if (tm) {
        float ms = 0.f;
        for (int64_t c = 0; c < nchunk; ++c) {
            T3CK(hipEventElapsedTime(&ms, s->ev[3 * c], s->ev[3 * c + 1]));
            tm->gemm_ms += ms;
            T3CK(hipEventElapsedTime(&ms, s->ev[3 * c + 1], s->ev[3 * c + 2]));
            tm->energy_ms += ms;
        }
        T3CK(hipEventElapsedTime(&ms, e_beg, e_end));
        tm->total_ms = ms;
        tm->chunks = nchunk;
        tm->chunk = s->zmax;
    }
    return 0;
}

}  // namespace rcc
