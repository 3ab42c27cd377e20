// CPU emulator of the symmetry-blocked plan (test infrastructure): runs the
// operation program of molfdir_amd/csrc/rcc_blocked_plan.h on host arrays with
// plain loops, so the host-side plan (index maps, product descriptors, K
// splits, the term sequence) is checked against the reference fixtures
// without a GPU.  Built by tests/test_blocked_plan.py with g++.
#include <complex>
#include <cstring>
#include <cstdio>
#include "rcc_blocked_plan.h"
using namespace rcc::blk;
typedef std::complex<double> C;
struct Emu {
    Plan P;
    std::vector<std::vector<C>> a;   // per array
    std::vector<std::vector<double>> r;
};
static C* ptr(Emu* e, int id) {
    if (id == A_T1) return e->a[A_T].data();
    if (id == A_T2) return e->a[A_T].data() + e->P.arr[A_T1].size;
    if (id == A_S1) return e->a[A_S].data();
    if (id == A_S2) return e->a[A_S].data() + e->P.arr[A_S1].size;
    return e->a[id].data();
}
extern "C" {
Emu* emu_create(int nrep, const int* no, const int* nv, const int* ptab, const int* frozen) {
    SymIn s;
    s.nrep = nrep;
    s.no.assign(no, no + nrep);
    s.nv.assign(nv, nv + nrep);
    s.ptab.assign(ptab, ptab + nrep * nrep);
    int O = 0;
    for (int i = 0; i < nrep; ++i) O += no[i];
    if (frozen) s.frozen.assign(frozen, frozen + O);
    Emu* e = new Emu();
    if (build_plan(s, e->P) != 0) { fprintf(stderr, "plan: %s\n", e->P.err.c_str()); return nullptr; }
    e->a.assign(A_COUNT, {});
    e->r.assign(A_COUNT, {});
    for (int i = 0; i < A_COUNT; ++i) {
        if (i == A_T1 || i == A_T2 || i == A_S1 || i == A_S2) continue;
        int64_t n = i == A_WS ? e->P.ws_elems : e->P.arr[i].size;
        e->a[i].assign(std::max<int64_t>(n, 1), C(0, 0));
    }
    return e;
}
int64_t emu_size(Emu* e, int id) { return e->P.arr[id].size; }
void emu_set(Emu* e, int id, const double* x) { memcpy(ptr(e, id), x, e->P.arr[id].size * 16); }
void emu_get(Emu* e, int id, double* x) { memcpy(x, ptr(e, id), e->P.arr[id].size * 16); }
void emu_set_eps(Emu* e, const double* eps) {
    std::vector<double> d1, d2, f1, f2;
    build_denominators(e->P, eps, d1, d2, f1, f2);
    e->r[A_D1] = d1; e->r[A_D2] = d2; e->r[A_F1m] = f1; e->r[A_F2m] = f2;
}
int emu_run(Emu* e, int prog) {
    Plan& P = e->P;
    for (const Op& op : P.prog[prog]) {
        if (op.kind == OK_GATHER) {
            const GatherOp& G = P.gathers[op.idx];
            C* out = ptr(e, G.out);
            for (int64_t x = 0; x < G.n; ++x) {
                C acc(0, 0);
                for (int t = 0; t < G.T; ++t) {
                    int32_t m = G.map[(size_t)t * G.n + x];
                    if (m < 0) continue;
                    C v = ptr(e, G.src[G.slot[t]])[m >> 2];
                    if (m & 1) v = std::conj(v);
                    if (m & 2) v = -v;
                    acc += v;
                }
                out[x] = (G.beta != 0.0 ? G.beta * out[x] : C(0, 0)) + G.alpha * acc;
            }
        } else if (op.kind == OK_BIL) {
            const BilOp& B = P.bils[op.idx];
            C* out = ptr(e, B.out);
            const C* s0 = ptr(e, B.src0); const C* A = ptr(e, B.A); const C* Bm = ptr(e, B.B);
            for (int64_t x = 0; x < B.n; ++x) {
                C v(0, 0);
                int32_t m = B.m0[x];
                if (m >= 0) { C w = s0[m >> 2]; if (m & 1) w = std::conj(w); if (m & 2) w = -w; v = B.c0 * w; }
                for (int t = 0; t < B.np; ++t)
                    if (B.p[t][x] >= 0) v += B.sg[t] * (A[B.p[t][x]] * Bm[B.q[t][x]]);
                out[x] = v;
            }
        } else if (op.kind == OK_EW) {
            const EwOp& w = P.ews[op.idx];
            int64_t n = P.arr[w.y].size;
            C* y = ptr(e, w.y); const C* x = ptr(e, w.x);
            for (int64_t i = 0; i < n; ++i) {
                if (w.kind == EW_COPY) y[i] = x[i];
                else if (w.kind == EW_CONJ) y[i] = std::conj(x[i]);
                else if (w.kind == EW_AXPBY) y[i] = (w.b != 0 ? w.b * y[i] : C(0,0)) + w.a * x[i];
                else y[i] = 0;
            }
        } else if (op.kind == OK_DENOM) {
            const DenomOp& d = P.denoms[op.idx];
            C* s1 = ptr(e, d.S1); C* s2 = ptr(e, d.S2);
            for (int64_t i = 0; i < P.arr[A_D1].size; ++i) {
                if (d.divide) s1[i] /= e->r[A_D1][i];
                if (d.mask && P.any_frozen && e->r[A_F1m][i] != 0) s1[i] = 0;
            }
            for (int64_t i = 0; i < P.arr[A_D2].size; ++i) {
                if (d.divide) s2[i] /= e->r[A_D2][i];
                if (d.mask && P.any_frozen && e->r[A_F2m][i] != 0) s2[i] = 0;
            }
        } else {
            const ProdOp& p = P.prods[op.idx];
            for (const PGroup& g : p.g) {
                const C* A = ptr(e, g.A) + g.offA; const C* B = ptr(e, g.B) + g.offB; C* Cc = ptr(e, g.C) + g.offC;
                for (int j = 0; j < g.n; ++j)
                    for (int i = 0; i < g.m; ++i) {
                        C acc(0, 0);
                        for (int k = 0; k < g.k; ++k) {
                            C av = (p.opa == 0 || p.opa == 3) ? A[i + g.lda * (int64_t)k] : A[k + g.lda * (int64_t)i];
                            if (p.opa == 2 || p.opa == 3) av = std::conj(av);
                            C bv = (p.opb == 0 || p.opb == 3) ? B[k + g.ldb * (int64_t)j] : B[j + g.ldb * (int64_t)k];
                            if (p.opb == 2 || p.opb == 3) bv = std::conj(bv);
                            acc += av * bv;
                        }
                        C& cc = Cc[i + g.ldc * (int64_t)j];
                        cc = g.alpha * acc + (g.beta != 0 ? g.beta * cc : C(0, 0));
                    }
            }
            for (const RedDesc& r : p.red) {
                const C* W = e->a[A_WS].data() + r.offW; C* Cc = ptr(e, r.C) + r.offC;
                for (int j = 0; j < r.n; ++j)
                    for (int i = 0; i < r.m; ++i) {
                        C s(0, 0);
                        for (int q = 0; q < r.nsplit; ++q) s += W[(int64_t)q * r.m * r.n + i + (int64_t)r.m * j];
                        C& cc = Cc[i + r.ldc * (int64_t)j];
                        cc = r.alpha * s + (r.beta != 0 ? r.beta * cc : C(0, 0));
                    }
            }
        }
    }
    return 0;
}
int emu_nprods(Emu* e) { return (int)e->P.prods.size(); }
void emu_prod_info(Emu* e, int i, double* info) {
    const ProdOp& p = e->P.prods[i];
    info[0] = p.g.size(); info[1] = p.red.size(); info[2] = p.variant; info[3] = p.flops; info[4] = p.ws_elems;
}
}
extern "C" int emu_arr(Emu* e, const char* name) {
    for (int i = 0; i < A_COUNT; ++i)
        if (e->P.arr[i].name == name) return i;
    return -1;
}
extern "C" int emu_prog(const char* name) {
    static const char* names[P_COUNT] = {"TAU", "RING", "X", "HOINTM", "HVINTM", "HOVINTM",
        "T1EQNS", "GOINTM", "GVINTM", "AINTM", "HINTM_HAKCI", "HINTM", "T2EQNS", "BINTM",
        "DENOM", "ZCORE", "TAUS", "MP2", "CONST", "DIVIDE"};
    for (int i = 0; i < P_COUNT; ++i)
        if (std::string(names[i]) == name) return i;
    return -1;
}
