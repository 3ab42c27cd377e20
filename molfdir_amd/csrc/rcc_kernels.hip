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
#include "rcc_internal.h"

namespace rcc {

constexpr int TPB = 256;

static inline int nblocks(int64_t n, int64_t cap = 1 << 20) {
    int64_t b = (n + TPB - 1) / TPB;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, cap));
}

__device__ __forceinline__ int64_t pair_of(int x, int y, int n) {
    // x > y
    return (int64_t)y * (2 * n - y - 1) / 2 + (x - y - 1);
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

// ---------------------------------------------------------------- tau
__global__ void gettau_kernel(const double2* __restrict__ t1, const double2* __restrict__ t2,
                              double2* __restrict__ tau, double s2, const int2* __restrict__ vv,
                              int64_t nvvt, const int2* __restrict__ oo, int64_t noot, int nv) {
    GRID_LOOP(e, nvvt * noot) {
        const int64_t ab = e % nvvt, ij = e / nvvt;
        const int2 p = vv[ab], q = oo[ij];
        const int a = p.x, b = p.y, i = q.x, j = q.y;
        double2 v = cscale(s2, t2[e]);
        v = cadd(v, cmul(t1[a + nv * i], t1[b + nv * j]));
        v = csub(v, cmul(t1[a + nv * j], t1[b + nv * i]));
        tau[e] = v;
    }
}

void k_gettau(const double2* t1, const double2* t2, double2* tau, double s2, PairTab vv,
              PairTab oo, int nv, hipStream_t st) {
    const int64_t n = vv.npair * oo.npair;
    if (n == 0) return;
    hipLaunchKernelGGL(gettau_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, t1, t2, tau, s2,
                       vv.tab, vv.npair, oo.tab, oo.npair, nv);
}

// ---------------------------------------------------------------- expansions
__global__ void expand_e2_kernel(const double2* __restrict__ P, double2* __restrict__ out,
                                 int64_t R, int no) {
    GRID_LOOP(o, R * no * no) {
        const int64_t r = o % R;
        const int64_t q = o / R;
        const int y = (int)(q % no), x = (int)(q / no);
        double2 v = make_double2(0.0, 0.0);
        if (x > y) v = P[r + R * pair_of(x, y, no)];
        else if (x < y) { v = P[r + R * pair_of(y, x, no)]; v.x = -v.x; v.y = -v.y; }
        out[o] = v;
    }
}

void k_expand_e2(const double2* P, double2* out, int64_t R, PairTab oo, int no, hipStream_t st) {
    const int64_t n = R * no * no;
    if (n == 0) return;
    hipLaunchKernelGGL(expand_e2_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, P, out, R, no);
}

__global__ void expand_f1_kernel(const double2* __restrict__ P, double2* __restrict__ out,
                                 int64_t ncol, int nv, int64_t nvvt) {
    const int64_t nv2 = (int64_t)nv * nv;
    GRID_LOOP(o, nv2 * ncol) {
        const int a = (int)(o % nv);
        const int d = (int)((o / nv) % nv);
        const int64_t col = o / nv2;
        double2 v = make_double2(0.0, 0.0);
        if (a > d) v = P[pair_of(a, d, nv) + nvvt * col];
        else if (a < d) { v = P[pair_of(d, a, nv) + nvvt * col]; v.x = -v.x; v.y = -v.y; }
        out[o] = v;
    }
}

void k_expand_f1(const double2* P, double2* out, int64_t ncol, PairTab vv, int nv,
                 hipStream_t st) {
    const int64_t n = (int64_t)nv * nv * ncol;
    if (n == 0) return;
    hipLaunchKernelGGL(expand_f1_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, P, out, ncol, nv,
                       vv.npair);
}

__global__ void expand_ring_kernel(const double2* __restrict__ T2, const double2* __restrict__ t1,
                                   double s, double2* __restrict__ out, int nv, int no,
                                   int64_t nvvt) {
    const int64_t nvo = (int64_t)nv * no;
    GRID_LOOP(o, nvo * nvo) {
        const int a = (int)(o % nv);
        const int i = (int)((o / nv) % no);
        const int64_t cc = o / nvo;
        const int c = (int)(cc % nv);
        const int k = (int)(cc / nv);
        double2 v = make_double2(0.0, 0.0);
        if (a != c && i != k) {
            const int64_t ab = a > c ? pair_of(a, c, nv) : pair_of(c, a, nv);
            const int64_t ij = i > k ? pair_of(i, k, no) : pair_of(k, i, no);
            v = T2[ab + nvvt * ij];
            const double sg = ((a > c) == (i > k)) ? s : -s;
            v = cscale(sg, v);
        }
        if (t1) v = csub(v, cmul(t1[a + nv * k], t1[c + nv * i]));
        out[o] = v;
    }
}

void k_expand_ring(const double2* T2, const double2* t1, double s, double2* out, PairTab vv,
                   PairTab oo, int nv, int no, hipStream_t st) {
    const int64_t nvo = (int64_t)nv * no;
    if (nvo == 0) return;
    hipLaunchKernelGGL(expand_ring_kernel, dim3(nblocks(nvo * nvo)), dim3(TPB), 0, st, T2, t1,
                       s, out, nv, no, vv.npair);
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
    const int64_t n = R * nv * nv;
    if (n == 0) return;
    hipLaunchKernelGGL(expand_cols_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, P, out, R, nv);
}

// ---------------------------------------------------------------- antisymmetrising packs
__global__ void pack_pij_kernel(const double2* __restrict__ Y, double2* __restrict__ S,
                                int64_t NR, double alpha, const int2* __restrict__ oo,
                                int64_t noot, int no) {
    GRID_LOOP(e, NR * noot) {
        const int64_t r = e % NR, ij = e / NR;
        const int2 q = oo[ij];
        const double2 y1 = Y[r + NR * (q.x + (int64_t)no * q.y)];
        const double2 y2 = Y[r + NR * (q.y + (int64_t)no * q.x)];
        double2 s = S[e];
        s.x += alpha * (y1.x - y2.x);
        s.y += alpha * (y1.y - y2.y);
        S[e] = s;
    }
}

void k_pack_pij(const double2* Y, double2* S, int64_t NR, double alpha, PairTab oo, int no,
                hipStream_t st) {
    const int64_t n = NR * oo.npair;
    if (n == 0) return;
    hipLaunchKernelGGL(pack_pij_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, Y, S, NR, alpha,
                       oo.tab, oo.npair, no);
}

__global__ void pack_pab_kernel(const double2* __restrict__ Y, double2* __restrict__ S,
                                int64_t ncol, double alpha, const int2* __restrict__ vv,
                                int64_t nvvt, int nv) {
    const int64_t nv2 = (int64_t)nv * nv;
    GRID_LOOP(e, nvvt * ncol) {
        const int64_t ab = e % nvvt, col = e / nvvt;
        const int2 p = vv[ab];
        const double2 y1 = Y[p.x + (int64_t)nv * p.y + nv2 * col];
        const double2 y2 = Y[p.y + (int64_t)nv * p.x + nv2 * col];
        double2 s = S[e];
        s.x += alpha * (y1.x - y2.x);
        s.y += alpha * (y1.y - y2.y);
        S[e] = s;
    }
}

void k_pack_pab(const double2* Y, double2* S, int64_t ncol, double alpha, PairTab vv, int nv,
                hipStream_t st) {
    const int64_t n = vv.npair * ncol;
    if (n == 0) return;
    hipLaunchKernelGGL(pack_pab_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, Y, S, ncol, alpha,
                       vv.tab, vv.npair, nv);
}

__global__ void pack_pabij_kernel(const double2* __restrict__ Q, double2* __restrict__ S,
                                  const int2* __restrict__ vv, int64_t nvvt,
                                  const int2* __restrict__ oo, int64_t noot, int nv, int no,
                                  double alpha) {
    const int64_t nvo = (int64_t)nv * no;
    GRID_LOOP(e, nvvt * noot) {
        const int2 p = vv[e % nvvt], q = oo[e / nvvt];
        const int64_t ai = p.x + (int64_t)nv * q.x, bj = p.y + (int64_t)nv * q.y;
        const int64_t bi = p.y + (int64_t)nv * q.x, aj = p.x + (int64_t)nv * q.y;
        const double2 v1 = Q[ai + nvo * bj], v2 = Q[bi + nvo * aj];
        const double2 v3 = Q[aj + nvo * bi], v4 = Q[bj + nvo * ai];
        double2 s = S[e];
        s.x += alpha * (v1.x - v2.x - v3.x + v4.x);
        s.y += alpha * (v1.y - v2.y - v3.y + v4.y);
        S[e] = s;
    }
}

void k_pack_pabij(const double2* Q, double2* S, PairTab vv, PairTab oo, int nv, int no,
                  hipStream_t st, double alpha) {
    const int64_t n = vv.npair * oo.npair;
    if (n == 0) return;
    hipLaunchKernelGGL(pack_pabij_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, Q, S, vv.tab,
                       vv.npair, oo.tab, oo.npair, nv, no, alpha);
}

// ---------------------------------------------------------------- ring permutation
__global__ void perm_akci_kernel(const double2* __restrict__ X, double2* __restrict__ H, int nv,
                                 int no, int accumulate) {
    const int64_t nvo = (int64_t)nv * no;
    GRID_LOOP(o, nvo * nvo) {
        const int a = (int)(o % nv);
        const int i = (int)((o / nv) % no);
        const int64_t cc = o / nvo;
        const int c = (int)(cc % nv), k = (int)(cc / nv);
        const double2 x = X[(a + (int64_t)nv * k) + nvo * (c + (int64_t)nv * i)];
        H[o] = accumulate ? cadd(H[o], x) : x;
    }
}

void k_perm_akci(const double2* X, double2* H, int nv, int no, bool accumulate,
                 hipStream_t st) {
    const int64_t nvo = (int64_t)nv * no;
    if (nvo == 0) return;
    hipLaunchKernelGGL(perm_akci_kernel, dim3(nblocks(nvo * nvo)), dim3(TPB), 0, st, X, H, nv,
                       no, accumulate ? 1 : 0);
}

// ---------------------------------------------------------------- traces
__global__ void trace_t1eqn2_kernel(const double2* __restrict__ X, double2* __restrict__ S1,
                                    int nv, int no) {
    const int64_t nvo = (int64_t)nv * no;
    GRID_LOOP(e, nvo) {
        const int a = (int)(e % nv), i = (int)(e / nv);
        double2 s = S1[e];
        for (int k = 0; k < no; ++k) {
            if (k == i) continue;
            const double2 x = X[(a + (int64_t)nv * k) + nvo * (i > k ? pair_of(i, k, no) : pair_of(k, i, no))];
            if (i > k) { s.x += x.x; s.y += x.y; }
            else { s.x -= x.x; s.y -= x.y; }
        }
        S1[e] = s;
    }
}

void k_trace_t1eqn2(const double2* X, double2* S1, PairTab oo, int nv, int no, hipStream_t st) {
    (void)oo;
    const int64_t n = (int64_t)nv * no;
    if (n == 0) return;
    hipLaunchKernelGGL(trace_t1eqn2_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, X, S1, nv, no);
}

__global__ void trace_gv_kernel(const double2* __restrict__ H1, double2* __restrict__ GV,
                                double alpha, int nv, int no) {
    const int64_t nvo = (int64_t)nv * no;
    GRID_LOOP(e, (int64_t)nv * nv) {
        const int a = (int)(e % nv), c = (int)(e / nv);
        double sr = 0.0, si = 0.0;
        for (int k = 0; k < no; ++k) {
            const double2 x = H1[(a + (int64_t)nv * k) + nvo * (c + (int64_t)nv * k)];
            sr += x.x;
            si += x.y;
        }
        double2 g = GV[e];
        g.x += alpha * sr;
        g.y += alpha * si;
        GV[e] = g;
    }
}

void k_trace_gv(const double2* H1, double2* GV, double alpha, int nv, int no, hipStream_t st) {
    const int64_t n = (int64_t)nv * nv;
    if (n == 0) return;
    hipLaunchKernelGGL(trace_gv_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, H1, GV, alpha, nv,
                       no);
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
__global__ void denom_t1_kernel(double2* __restrict__ S1, const double* __restrict__ eps,
                                const int* __restrict__ frozen, int nv, int no) {
    GRID_LOOP(e, (int64_t)nv * no) {
        const int a = (int)(e % nv), i = (int)(e / nv);
        double2 s = S1[e];
        if (eps) {
            const double d = eps[i] - eps[no + a];
            s.x /= d;
            s.y /= d;
        }
        if (frozen && frozen[i]) s = make_double2(0.0, 0.0);
        S1[e] = s;
    }
}

__global__ void denom_t2_kernel(double2* __restrict__ S2, const double* __restrict__ eps,
                                const int* __restrict__ frozen, const int2* __restrict__ vv,
                                int64_t nvvt, const int2* __restrict__ oo, int64_t noot, int no) {
    GRID_LOOP(e, nvvt * noot) {
        const int2 p = vv[e % nvvt], q = oo[e / nvvt];
        double2 s = S2[e];
        if (eps) {
            const double d = eps[q.x] + eps[q.y] - eps[no + p.x] - eps[no + p.y];
            s.x /= d;
            s.y /= d;
        }
        if (frozen && (frozen[q.x] || frozen[q.y])) s = make_double2(0.0, 0.0);
        S2[e] = s;
    }
}

void k_denom(double2* S1, double2* S2, const double* eps, const int* frozen, PairTab vv,
             PairTab oo, int nv, int no, hipStream_t st) {
    if (S1 && (int64_t)nv * no > 0)
        hipLaunchKernelGGL(denom_t1_kernel, dim3(nblocks((int64_t)nv * no)), dim3(TPB), 0, st,
                           S1, eps, frozen, nv, no);
    const int64_t n2 = vv.npair * oo.npair;
    if (S2 && n2 > 0)
        hipLaunchKernelGGL(denom_t2_kernel, dim3(nblocks(n2)), dim3(TPB), 0, st, S2, eps, frozen,
                           vv.tab, vv.npair, oo.tab, oo.npair, no);
}

// ---------------------------------------------------------------- small BLAS-1
__global__ void axpby_kernel(int64_t n, double2 a, const double2* __restrict__ x, double2 b,
                             double2* __restrict__ y) {
    GRID_LOOP(e, n) {
        double2 r = make_double2(0.0, 0.0);
        if (b.x != 0.0 || b.y != 0.0) r = cmul(b, y[e]);
        if (x) r = cadd(r, cmul(a, x[e]));
        y[e] = r;
    }
}

void k_axpby(int64_t n, double2 a, const double2* x, double2 b, double2* y, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(axpby_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, n, a, x, b, y);
}

__global__ void conj_kernel(int64_t n, const double2* __restrict__ x, double2* __restrict__ y) {
    GRID_LOOP(e, n) { y[e] = make_double2(x[e].x, -x[e].y); }
}

void k_conj(int64_t n, const double2* x, double2* y, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(conj_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, n, x, y);
}

__global__ void zero_diag_kernel(double2* M, int n) {
    GRID_LOOP(e, n) { M[e + (int64_t)n * e] = make_double2(0.0, 0.0); }
}

void k_zero_diag(double2* M, int n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(zero_diag_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, M, n);
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

// Block b reduces elements [b*chunk, min(n,(b+1)*chunk)) of sum conj(x_v) * y for each v.
__global__ void __launch_bounds__(TPB) zdotc_part_kernel(PtrPack xs, int nx,
                                                         const double2* __restrict__ y, int64_t n,
                                                         int64_t chunk, double2* __restrict__ part,
                                                         int nblk) {
    __shared__ double sh[2][TPB];
    const int64_t beg = blockIdx.x * chunk;
    const int64_t end = std::min<int64_t>(n, beg + chunk);
    for (int v = 0; v < nx; ++v) {
        const double2* __restrict__ x = xs.p[v];
        double sr = 0.0, si = 0.0;
        for (int64_t e = beg + threadIdx.x; e < end; e += TPB) {
            const double2 a = x[e], b = y[e];
            sr += a.x * b.x + a.y * b.y;
            si += a.x * b.y - a.y * b.x;
        }
        sh[0][threadIdx.x] = sr;
        sh[1][threadIdx.x] = si;
        __syncthreads();
        for (int s = TPB / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
                sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) part[v * nblk + blockIdx.x] = make_double2(sh[0][0], sh[1][0]);
        __syncthreads();
    }
}

// One workgroup per reduced value: strided partial sums then a tree, both in a
// fixed order, so the result does not depend on scheduling.
__global__ void __launch_bounds__(TPB) finish_kernel(const double2* __restrict__ part, int nblk,
                                                     int nv, double2* __restrict__ dst) {
    __shared__ double sh[2][TPB];
    const int v = blockIdx.x;
    double sr = 0.0, si = 0.0;
    for (int b = threadIdx.x; b < nblk; b += TPB) {
        sr += part[v * nblk + b].x;
        si += part[v * nblk + b].y;
    }
    sh[0][threadIdx.x] = sr;
    sh[1][threadIdx.x] = si;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) dst[v] = make_double2(sh[0][0], sh[1][0]);
}

// partial-sum blocks for n elements: about 4096 elements each, at most nblk
static inline int red_blocks(int64_t n, int nblk) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(nblk, (n + 4095) / 4096));
}

void k_zdotc_multi(const double2* const* xs, int nx, const double2* y, int64_t n, double2* dst,
                   double2* part, int nblk, hipStream_t st) {
    if (nx <= 0) return;
    PtrPack pk{};
    for (int v = 0; v < nx && v < MAXV; ++v) pk.p[v] = xs[v];
    nblk = red_blocks(n, nblk);
    const int64_t chunk = (n + nblk - 1) / nblk;
    hipLaunchKernelGGL(zdotc_part_kernel, dim3(nblk), dim3(TPB), 0, st, pk, nx, y, n,
                       std::max<int64_t>(chunk, 1), part, nblk);
    hipLaunchKernelGGL(finish_kernel, dim3(nx), dim3(TPB), 0, st, part, nblk, nx, dst);
}

// dst[0] = sum x, dst[1] = (sum |x|^2, 0)
__global__ void __launch_bounds__(TPB) sumnorm_part_kernel(const double2* __restrict__ x,
                                                           int64_t n, int64_t chunk,
                                                           double2* __restrict__ part, int nblk) {
    __shared__ double sh[3][TPB];
    const int64_t beg = blockIdx.x * chunk;
    const int64_t end = std::min<int64_t>(n, beg + chunk);
    double sr = 0.0, si = 0.0, nn = 0.0;
    for (int64_t e = beg + threadIdx.x; e < end; e += TPB) {
        const double2 a = x[e];
        sr += a.x;
        si += a.y;
        nn += a.x * a.x + a.y * a.y;
    }
    sh[0][threadIdx.x] = sr;
    sh[1][threadIdx.x] = si;
    sh[2][threadIdx.x] = nn;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
            sh[2][threadIdx.x] += sh[2][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = make_double2(sh[0][0], sh[1][0]);
        part[nblk + blockIdx.x] = make_double2(sh[2][0], 0.0);
    }
}

void k_sum_and_norm(const double2* x, int64_t n, double2* dst, double2* part, int nblk,
                    hipStream_t st) {
    nblk = red_blocks(n, nblk);
    const int64_t chunk = (n + nblk - 1) / nblk;
    hipLaunchKernelGGL(sumnorm_part_kernel, dim3(nblk), dim3(TPB), 0, st, x, n,
                       std::max<int64_t>(chunk, 1), part, nblk);
    hipLaunchKernelGGL(finish_kernel, dim3(2), dim3(TPB), 0, st, part, nblk, 2, dst);
}

__global__ void lincomb_kernel(CoefPack cp, int nx, double2* __restrict__ y, int64_t n) {
    GRID_LOOP(e, n) {
        double sr = 0.0, si = 0.0;
        for (int v = 0; v < nx; ++v) {
            const double2 a = cp.p[v][e];
            sr += cp.c[v] * a.x;
            si += cp.c[v] * a.y;
        }
        y[e] = make_double2(sr, si);
    }
}

void k_lincomb(const double2* const* xs, const double* coef, int nx, double2* y, int64_t n,
               hipStream_t st) {
    CoefPack cp{};
    for (int v = 0; v < nx && v < MAXV; ++v) {
        cp.p[v] = xs[v];
        cp.c[v] = coef[v];
    }
    if (n <= 0) return;
    hipLaunchKernelGGL(lincomb_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, cp, nx, y, n);
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

__global__ void fill_zero_kernel(double2* __restrict__ p, int64_t n) {
    GRID_LOOP(e, n) { p[e] = make_double2(0.0, 0.0); }
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
__global__ void scatter_set_kernel(const double2* __restrict__ src, BlockSet bs,
                                   double2* __restrict__ dst, int64_t ld, int gather) {
    const BlockSet::Blk& b = bs.b[blockIdx.y];
    const int64_t n = b.rows * b.cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e % b.rows, c = e / b.rows;
        const int64_t c1 = (int64_t)b.rmap[r] + ld * b.cmap[c];
        if (gather) dst[b.off + e] = src[c1];
        else dst[c1] = src[b.off + e];
    }
}

void k_blockset(const double2* src, const BlockSet& bs, double2* dst, int64_t ld, bool gather,
                hipStream_t st) {
    int64_t mx = 0;
    for (int i = 0; i < bs.nb; ++i) mx = std::max<int64_t>(mx, bs.b[i].rows * bs.b[i].cols);
    if (bs.nb <= 0 || mx <= 0) return;
    hipLaunchKernelGGL(scatter_set_kernel, dim3(nblocks(mx), bs.nb), dim3(TPB), 0, st, src, bs,
                       dst, ld, gather ? 1 : 0);
}

void k_fill_zero(double2* p, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, p, n);
}

}  // namespace rcc
