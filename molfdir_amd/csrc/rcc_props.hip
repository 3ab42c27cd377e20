// Device work of the MP2 relaxed first-order property: the two integral x
// density contractions of the X intermediate (CCDENSZ ccgrad.F:199-221) and
// the zero-frequency Z-vector iteration of RPAA (ccrpa.F:34-282) with its
// matrix-vector products DOAX / DOBY (ccrpa.F:467-603), the preconditioner
// GETDG (ccrpa.F:1450-1528) and PRECD (ccrpa.F:695-754).
//
// All reductions run in a fixed order (per-column partial sums, then one
// thread per output; no atomics), so repeated calls give bit-identical
// results.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "rcc_internal.h"

namespace rcc {

namespace {

__device__ __forceinline__ int64_t pidx(int p, int q, int n) {
    return (int64_t)q * (2 * n - q - 1) / 2 + (p - q - 1);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// X[a + nv*i] += sum_{k,l} <ak||il> D[k + no*l] over the packed VOOO class
// V[(a + nv*k), pair(x > y)] = <ak||xy>, in two passes.  Pass 1, one
// workgroup per column pair(x > y), each element read once:
//   part[2 (a + nv P)]     = sum_k V[(a,k),P] D(k,y)   (-> X(a,x) +=)
//   part[2 (a + nv P) + 1] = sum_k V[(a,k),P] D(k,x)   (-> X(a,y) -=)
// Pass 2, one thread per (a, m): X(a,m) += sum_{y<m} part1[pair(m,y)] -
// sum_{x>m} part2[pair(x,m)], in a fixed order.
__global__ void __launch_bounds__(256)
x_vooo_cols_kernel(const double2* __restrict__ V, const double2* __restrict__ D,
                   double2* __restrict__ part, PairTab oo, int nv, int no) {
    const int64_t nvo = (int64_t)nv * no;
    const int64_t P = blockIdx.x;
    const int2 xy = oo.tab[P];
    const double2* col = V + nvo * P;
    for (int a = threadIdx.x; a < nv; a += blockDim.x) {
        double2 s1 = make_double2(0.0, 0.0), s2 = make_double2(0.0, 0.0);
        for (int k = 0; k < no; ++k) {
            const double2 v = col[a + (int64_t)nv * k];
            const double2 p1 = cmul(v, D[k + (int64_t)no * xy.y]);
            const double2 p2 = cmul(v, D[k + (int64_t)no * xy.x]);
            s1.x += p1.x; s1.y += p1.y;
            s2.x += p2.x; s2.y += p2.y;
        }
        part[2 * (a + (int64_t)nv * P)] = s1;
        part[2 * (a + (int64_t)nv * P) + 1] = s2;
    }
}

__global__ void x_vooo_finish_kernel(const double2* __restrict__ part, double2* __restrict__ X,
                                     int nv, int no) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nv * no) return;
    const int a = (int)(e % nv), m = (int)(e / nv);
    double re = 0.0, im = 0.0;
    for (int y = 0; y < m; ++y) {
        const double2 p = part[2 * (a + (int64_t)nv * pidx(m, y, no))];
        re += p.x;
        im += p.y;
    }
    for (int x = m + 1; x < no; ++x) {
        const double2 p = part[2 * (a + (int64_t)nv * pidx(x, m, no)) + 1];
        re -= p.x;
        im -= p.y;
    }
    X[e].x += re;
    X[e].y += im;
}

// X[a + nv*i] += sum_{b,c} D[b + nv*c] conj <bi||ca> over the packed VOVV
// class V[(b + nv*i), pair(x > y)] = <bi||xy>, in two passes.  Pass 1, one
// workgroup per column pair(x > y) (each element read once), one wave per
// occupied i at a time, lanes over b, fixed-order butterfly reduction:
//   part[2 (i + no P)]     = sum_b D(b,x) conj V[(b,i),P]   (-> X(y,i) +=)
//   part[2 (i + no P) + 1] = sum_b D(b,y) conj V[(b,i),P]   (-> X(x,i) -=)
// Pass 2, one thread per (m, i): X(m,i) += sum_{x>m} part1[pair(x,m)] -
// sum_{y<m} part2[pair(m,y)].
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ void __launch_bounds__(256)
x_vovv_cols_kernel(const double2* __restrict__ V, const double2* __restrict__ D,
                   double2* __restrict__ part, PairTab vv, int nv, int no) {
    const int64_t nvo = (int64_t)nv * no;
    const int64_t P = blockIdx.x;
    const int2 xy = vv.tab[P];
    const double2* col = V + nvo * P;
    const double2* dx = D + (int64_t)nv * xy.x;
    const double2* dy = D + (int64_t)nv * xy.y;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = w; i < no; i += 4) {
        double r1 = 0.0, i1 = 0.0, r2 = 0.0, i2 = 0.0;
        for (int b = lane; b < nv; b += 64) {
            const double2 v = col[b + (int64_t)nv * i];
            const double2 d1 = dx[b], d2 = dy[b];
            r1 += d1.x * v.x + d1.y * v.y;
            i1 += d1.y * v.x - d1.x * v.y;
            r2 += d2.x * v.x + d2.y * v.y;
            i2 += d2.y * v.x - d2.x * v.y;
        }
        r1 = wave_sum(r1);
        i1 = wave_sum(i1);
        r2 = wave_sum(r2);
        i2 = wave_sum(i2);
        if (lane == 0) {
            part[2 * (i + (int64_t)no * P)] = make_double2(r1, i1);
            part[2 * (i + (int64_t)no * P) + 1] = make_double2(r2, i2);
        }
    }
}

__global__ void x_vovv_finish_kernel(const double2* __restrict__ part, double2* __restrict__ X,
                                     int nv, int no) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nv * no) return;
    const int i = (int)(e % no), m = (int)(e / no);
    double re = 0.0, im = 0.0;
    for (int x = m + 1; x < nv; ++x) {
        const double2 p = part[2 * (i + (int64_t)no * pidx(x, m, nv))];
        re += p.x;
        im += p.y;
    }
    for (int y = 0; y < m; ++y) {
        const double2 p = part[2 * (i + (int64_t)no * pidx(m, y, nv)) + 1];
        re -= p.x;
        im -= p.y;
    }
    double2* xp = X + m + (int64_t)nv * i;
    xp->x += re;
    xp->y += im;
}

// dg[a + nv*i] = eps[no + a] - eps[i] - Re R[(ai), (ai)]   (GETDG)
__global__ void zvec_dg_kernel(const double2* __restrict__ R, const double* __restrict__ eps,
                               double* __restrict__ dg, int nv, int no) {
    const int64_t n = (int64_t)nv * no;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int a = (int)(e % nv), i = (int)(e / nv);
    dg[e] = (eps[no + a] - eps[i]) - R[e + n * e].x;
}

// part[ch*n + r] = sum over columns c of chunk ch of R[r + n*c] x[c] - B[r + n*c] conj x[c]
// (DOAX + DOBY with Y = conj X; both matrices read once, column-major,
// consecutive threads on consecutive rows)
constexpr int ZV_THREADS = 256;

__global__ void __launch_bounds__(ZV_THREADS)
zvec_gemv_kernel(const double2* __restrict__ R, const double2* __restrict__ B,
                 const double2* __restrict__ x, double2* __restrict__ part, int64_t n, int chunk) {
    const int64_t r = (int64_t)blockIdx.x * ZV_THREADS + threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.y * chunk;
    const int64_t c1 = min(n, c0 + chunk);
    if (r >= n) return;
    double re = 0.0, im = 0.0;
    for (int64_t c = c0; c < c1; ++c) {
        const double2 xc = x[c];
        const double2 a = R[r + n * c], b = B[r + n * c];
        // a x - b conj(x)
        re += (a.x * xc.x - a.y * xc.y) - (b.x * xc.x + b.y * xc.y);
        im += (a.x * xc.y + a.y * xc.x) - (b.y * xc.x - b.x * xc.y);
    }
    part[(int64_t)blockIdx.y * n + r] = make_double2(re, im);
}

// xn[r] = (rhs[r] + sum_ch part[ch*n + r] - R[r,r] x[r]) / dg[r]   (the
// diagonal of A is zeroed for the linear response, DOAX ccrpa.F:526-527)
__global__ void zvec_step_kernel(const double2* __restrict__ part, int nch,
                                 const double2* __restrict__ rhs, const double2* __restrict__ R,
                                 const double2* __restrict__ x, const double* __restrict__ dg,
                                 double2* __restrict__ xn, int64_t n) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double re = 0.0, im = 0.0;
    for (int ch = 0; ch < nch; ++ch) {
        const double2 p = part[(int64_t)ch * n + r];
        re += p.x;
        im += p.y;
    }
    const double2 d = cmul(R[r + n * r], x[r]);
    re = re - d.x + rhs[r].x;
    im = im - d.y + rhs[r].y;
    xn[r] = make_double2(re / dg[r], im / dg[r]);
}

// y = x / dg   (PRECD at zero frequency)
__global__ void zvec_precd_kernel(const double2* __restrict__ x, const double* __restrict__ dg,
                                  double2* __restrict__ y, int64_t n) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    y[r] = make_double2(x[r].x / dg[r], x[r].y / dg[r]);
}

inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void k_x_vooo_doo(const double2* vooo, const double2* doo, double2* x, PairTab oo, double2* part,
                  int nv, int no, hipStream_t st) {
    if (oo.npair == 0 || nv == 0) return;
    hipLaunchKernelGGL(x_vooo_cols_kernel, dim3((unsigned)oo.npair), dim3(256), 0, st, vooo, doo,
                       part, oo, nv, no);
    hipLaunchKernelGGL(x_vooo_finish_kernel, dim3(nblk((int64_t)nv * no, 256)), dim3(256), 0, st,
                       part, x, nv, no);
}

void k_x_vovv_dvv(const double2* vovv, const double2* dvv, double2* x, PairTab vv, double2* part,
                  int nv, int no, hipStream_t st) {
    if (vv.npair == 0 || no == 0) return;
    hipLaunchKernelGGL(x_vovv_cols_kernel, dim3((unsigned)vv.npair), dim3(256), 0, st, vovv, dvv,
                       part, vv, nv, no);
    hipLaunchKernelGGL(x_vovv_finish_kernel, dim3(nblk((int64_t)nv * no, 256)), dim3(256), 0, st,
                       part, x, nv, no);
}

void k_zvec_dg(const double2* R, const double* eps, double* dg, int nv, int no, hipStream_t st) {
    const int64_t n = (int64_t)nv * no;
    if (n == 0) return;
    hipLaunchKernelGGL(zvec_dg_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, R, eps, dg, nv, no);
}

int zvec_chunks(int64_t n) {
    // columns per chunk: enough workgroups to fill the chip at small n,
    // short partial-sum passes at large n
    const int64_t chunk = n <= 2048 ? 32 : 128;
    return (int)((n + chunk - 1) / chunk);
}

void k_zvec_gemv(const double2* R, const double2* B, const double2* x, double2* part, int64_t n,
                 hipStream_t st) {
    if (n == 0) return;
    const int nch = zvec_chunks(n);
    const int chunk = (int)((n + nch - 1) / nch);
    hipLaunchKernelGGL(zvec_gemv_kernel, dim3(nblk(n, ZV_THREADS), nch), dim3(ZV_THREADS), 0, st,
                       R, B, x, part, n, chunk);
}

void k_zvec_step(const double2* part, const double2* rhs, const double2* R, const double2* x,
                 const double* dg, double2* xn, int64_t n, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(zvec_step_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, part,
                       zvec_chunks(n), rhs, R, x, dg, xn, n);
}

void k_zvec_precd(const double2* x, const double* dg, double2* y, int64_t n, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(zvec_precd_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, x, dg, y, n);
}

}  // namespace rcc
