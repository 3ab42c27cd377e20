// Device work of the MP2 relaxed first-order property: the two integral x
// density contractions of the X intermediate (CCDENSZ ccgrad.F:199-221) and
// the zero-frequency Z-vector iteration of RPAA (ccrpa.F:34-282) with its
// matrix-vector products DOAX / DOBY (ccrpa.F:467-603), the preconditioner
// GETDG (ccrpa.F:1450-1528) and PRECD (ccrpa.F:695-754).
//
// All reductions run in a fixed order (one thread or one workgroup per
// output, no atomics), so repeated calls give bit-identical results.
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

// X[a + nv*i] += sum_{k,l} <ak||il> D[k + no*l]; V is the packed VOOO class
// [(a + nv*k), pair(p > q)] = <ak||pq>.  One thread per (a, i); consecutive
// threads read consecutive rows of a column.
__global__ void x_vooo_doo_kernel(const double2* __restrict__ V, const double2* __restrict__ D,
                                  double2* __restrict__ X, int nv, int no) {
    const int64_t nvo = (int64_t)nv * no;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nvo) return;
    const int a = (int)(e % nv), i = (int)(e / nv);
    double re = 0.0, im = 0.0;
    for (int l = 0; l < no; ++l) {
        if (l == i) continue;
        const int64_t col = nvo * (i > l ? pidx(i, l, no) : pidx(l, i, no));
        const double s = i > l ? 1.0 : -1.0;
        for (int k = 0; k < no; ++k) {
            const double2 p = cmul(V[col + a + (int64_t)nv * k], D[k + (int64_t)no * l]);
            re += s * p.x;
            im += s * p.y;
        }
    }
    X[e].x += re;
    X[e].y += im;
}

// X[a + nv*i] += sum_{b,c} D[b + nv*c] conj <bi||ca>; V is the packed VOVV
// class [(b + nv*i), pair(p > q)] = <bi||pq>.  Workgroup (a, block of IC
// occupied i): for every c != a it reads the IC*nv contiguous rows (b, i) of
// column pair(c, a), i.e. each element of the class is read twice in all
// (once for each index of its pair).  Fixed-order workgroup reduction.
constexpr int XV_IC = 8;
constexpr int XV_THREADS = 128;

__global__ void __launch_bounds__(XV_THREADS)
x_vovv_dvv_kernel(const double2* __restrict__ V, const double2* __restrict__ D,
                  double2* __restrict__ X, int nv, int no) {
    const int64_t nvo = (int64_t)nv * no;
    const int a = blockIdx.x;
    const int i0 = blockIdx.y * XV_IC;
    const int ni = min(XV_IC, no - i0);
    double re[XV_IC], im[XV_IC];
#pragma unroll
    for (int q = 0; q < XV_IC; ++q) re[q] = im[q] = 0.0;
    for (int c = 0; c < nv; ++c) {
        if (c == a) continue;
        const double2* col = V + nvo * (c > a ? pidx(c, a, nv) : pidx(a, c, nv)) + (int64_t)nv * i0;
        const double s = c > a ? 1.0 : -1.0;
        const double2* dc = D + (int64_t)nv * c;
        for (int b = threadIdx.x; b < nv; b += XV_THREADS) {
            const double2 d = dc[b];
            const double dr = s * d.x, di = s * d.y;
#pragma unroll
            for (int q = 0; q < XV_IC; ++q) {
                if (q < ni) {
                    const double2 v = col[b + (int64_t)nv * q];
                    // d * conj(v)
                    re[q] += dr * v.x + di * v.y;
                    im[q] += di * v.x - dr * v.y;
                }
            }
        }
    }
    __shared__ double sr[XV_IC][XV_THREADS], si[XV_IC][XV_THREADS];
#pragma unroll
    for (int q = 0; q < XV_IC; ++q) {
        sr[q][threadIdx.x] = re[q];
        si[q][threadIdx.x] = im[q];
    }
    __syncthreads();
    for (int w = XV_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int q = 0; q < XV_IC; ++q) {
                sr[q][threadIdx.x] += sr[q][threadIdx.x + w];
                si[q][threadIdx.x] += si[q][threadIdx.x + w];
            }
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < ni) {
        const int q = threadIdx.x;
        double2* x = X + a + (int64_t)nv * (i0 + q);
        x->x += sr[q][0];
        x->y += si[q][0];
    }
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

void k_x_vooo_doo(const double2* vooo, const double2* doo, double2* x, int nv, int no,
                  hipStream_t st) {
    const int64_t n = (int64_t)nv * no;
    if (n == 0 || no < 2) return;
    hipLaunchKernelGGL(x_vooo_doo_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, vooo, doo, x, nv,
                       no);
}

void k_x_vovv_dvv(const double2* vovv, const double2* dvv, double2* x, int nv, int no,
                  hipStream_t st) {
    if (nv < 2 || no < 1) return;
    hipLaunchKernelGGL(x_vovv_dvv_kernel, dim3(nv, (no + XV_IC - 1) / XV_IC), dim3(XV_THREADS), 0,
                       st, vovv, dvv, x, nv, no);
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
