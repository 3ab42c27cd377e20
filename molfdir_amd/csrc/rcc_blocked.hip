// Symmetry-blocked CCSD iteration on the device: runs the operation program
// of rcc_blocked_plan.h (grouped products over the irrep blocks, map gathers,
// tau and the ring intermediate, denominators).  Every array of the program
// holds only the nonzero blocks of its symmetry, so the work and the memory
// of the iteration follow the blocks, as relccsd's per-irrep CNTRCT calls do
// (ccintm.F:1662-1723), instead of the C1 sizes.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>
#include "rcc_internal.h"
#include "rcc_blocked.h"
#include "rcc_blocked_plan.h"

namespace rcc {

using namespace blk;

struct BlkEngine {
    Plan P;
    std::vector<double2*> base;          // per array (T1 / T2 / S1 / S2: inside T / S)
    std::vector<bool> owned;
    std::vector<void*> allocs;
    std::vector<int32_t*> gmap;          // per gather
    struct BilDev { int32_t *m0 = nullptr, *p[2] = {nullptr, nullptr}, *q[2] = {nullptr, nullptr}; };
    std::vector<BilDev> bdev;
    struct ProdDev {
        std::vector<ZGroup> h96;         // host descriptors (pointers resolved)
        ZGroup* dg = nullptr;            // device copy, offsets for the product's kernel
        int64_t blocks = 0;
        GRedDesc* red = nullptr;
        int nred = 0;
        int64_t redmax = 0;
    };
    std::vector<ProdDev> pdev;
    double *D1 = nullptr, *D2 = nullptr, *F1m = nullptr, *F2m = nullptr;
    bool eps_ok = false;
    bool tau_ok = false, ring_ok = false, x_ok = false;
    int64_t bytes = 0;
};

namespace {

template <class T>
int balloc(BlkEngine* E, T** p, int64_t bytes, std::string& err) {
    *p = nullptr;
    const hipError_t e = hipMalloc((void**)p, (size_t)std::max<int64_t>(bytes, 16));
    if (e != hipSuccess) {
        err = "blocked engine: hipMalloc of " + std::to_string((double)bytes / 1e9) +
              " GB failed: " + hipGetErrorString(e);
        (void)hipGetLastError();
        *p = nullptr;
        return -1;
    }
    E->allocs.push_back((void*)*p);
    E->bytes += bytes;
    return 0;
}

template <class T>
int upload_vec(BlkEngine* E, T** dst, const std::vector<T>& v, std::string& err) {
    if (balloc(E, dst, (int64_t)v.size() * (int64_t)sizeof(T), err) != 0) return -1;
    if (!v.empty() &&
        hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
        err = "blocked engine: upload failed";
        return -1;
    }
    return 0;
}

ZGroup zg(const BlkEngine* E, const PGroup& g) {
    ZGroup z{};
    z.m = g.m;
    z.n = g.n;
    z.k = g.k;
    z.A = E->base[g.A] + g.offA;
    z.lda = g.lda;
    z.B = E->base[g.B] + g.offB;
    z.ldb = g.ldb;
    z.C = E->base[g.C] + g.offC;
    z.ldc = g.ldc;
    z.alpha = make_double2(g.alpha, 0.0);
    z.beta = make_double2(g.beta, 0.0);
    return z;
}

// descriptor tables of every product (after all arrays are bound)
int build_products(BlkEngine* E, std::string& err) {
    const Plan& P = E->P;
    E->pdev.assign(P.prods.size(), BlkEngine::ProdDev());
    for (size_t i = 0; i < P.prods.size(); ++i) {
        const ProdOp& op = P.prods[i];
        BlkEngine::ProdDev& d = E->pdev[i];
        for (const PGroup& g : op.g) d.h96.push_back(zg(E, g));
        if (!d.h96.empty()) {
            std::vector<ZGroup> a = d.h96;
            const int ng = (int)a.size();
            d.blocks = op.variant == 1   ? zgroup_offsets_small(a.data(), ng)
                       : op.variant == 2 ? zgroup_offsets_tall(a.data(), ng)
                       : op.variant == 3 ? zgroup_offsets_skinny(a.data(), ng)
                                         : zgroup_offsets(a.data(), ng);
            if (upload_vec(E, &d.dg, a, err) != 0) return -1;
        }
        if (!op.red.empty()) {
            std::vector<GRedDesc> r;
            for (const RedDesc& q : op.red) {
                GRedDesc g{};
                g.C = E->base[q.C] + q.offC;
                g.W = E->base[A_WS] + q.offW;
                g.ldc = q.ldc;
                g.m = q.m;
                g.n = q.n;
                g.nsplit = q.nsplit;
                g.alpha = q.alpha;
                g.beta = q.beta;
                r.push_back(g);
                d.redmax = std::max<int64_t>(d.redmax, (int64_t)q.m * q.n * split_ways(q.nsplit));
            }
            d.nred = (int)r.size();
            if (upload_vec(E, &d.red, r, err) != 0) return -1;
        }
    }
    return 0;
}

int run_prod(BlkEngine* E, int i, int gauss, hipStream_t st, std::string& err) {
    const ProdOp& op = E->P.prods[i];
    BlkEngine::ProdDev& d = E->pdev[i];
    // the same launches for every handle, whatever its blocks (a recorded
    // sequence must not depend on the sizes): empty products and reductions
    // are recorded too
    if (g_record || !d.h96.empty()) {
        if (g_record) {
            Rec r;
            r.kind = RK_GROUPED;
            r.groups = d.h96;
            r.opa = op.opa;
            r.opb = op.opb;
            r.tag = op.tag;
            r.gauss = gauss;
            if (op.variant == 3) {
                r.variant = 3;
                r.nj = op.nj;
            }
            g_record->push_back(std::move(r));
        } else {
            const int ng = (int)d.h96.size();
            const int rc =
                op.variant == 1   ? zgemm_grouped_small(op.opa, op.opb, gauss, d.dg, ng, d.blocks,
                                                        st, nullptr, 0)
                : op.variant == 2 ? zgemm_grouped_tall(op.opa, op.opb, gauss, d.dg, ng, d.blocks, st)
                : op.variant == 3 ? zgemm_grouped_skinny(op.nj, d.dg, ng, d.blocks, st)
                                  : zgemm_grouped_ops(op.opa, op.opb, op.tag, gauss, d.dg, ng,
                                                      d.blocks, st);
            if (rc != 0) {
                err = "blocked engine: grouped product " + op.name + " failed (" +
                      std::to_string(rc) + ")";
                return -1;
            }
        }
    }
    k_gred(d.red, d.nred, d.redmax, st);
    return 0;
}

int run_gather(BlkEngine* E, int i, hipStream_t st) {
    const GatherOp& G = E->P.gathers[i];
    const double2* src[4] = {nullptr, nullptr, nullptr, nullptr};
    for (size_t k = 0; k < G.src.size(); ++k) src[k] = E->base[G.src[k]];
    int slot[4] = {0, 0, 0, 0};
    for (int t = 0; t < G.T && t < 4; ++t) slot[t] = G.slot[t];
    k_map_sum(E->base[G.out], src, slot, (int)G.src.size(), E->gmap[i], G.n, G.T, G.alpha,
              G.beta, st);
    return 0;
}

void bil_args(const BlkEngine* E, int i, void* q) {
    const BilOp& B = E->P.bils[i];
    const BlkEngine::BilDev& d = E->bdev[i];
    mapbil_params(E->base[B.out], E->base[B.src0], E->base[B.A], E->base[B.B], d.m0, d.p[0],
                  d.q[0], d.p[1], d.q[1], B.n, B.np, B.c0, B.sg[0], B.sg[1], q);
}

int run_bil(BlkEngine* E, int i, hipStream_t st) {
    const BilOp& B = E->P.bils[i];
    const BlkEngine::BilDev& d = E->bdev[i];
    k_map_bil(E->base[B.out], E->base[B.src0], E->base[B.A], E->base[B.B], d.m0, d.p[0], d.q[0],
              d.p[1], d.q[1], B.n, B.np, B.c0, B.sg[0], B.sg[1], st);
    return 0;
}

int run_ew(BlkEngine* E, int i, hipStream_t st) {
    const EwOp& w = E->P.ews[i];
    const int64_t n = E->P.arr[w.y].size;
    double2* y = E->base[w.y];
    const double2* x = w.x >= 0 ? E->base[w.x] : nullptr;
    switch (w.kind) {
        case EW_COPY: k_copy(n, x, y, st); break;
        case EW_CONJ: k_conj(n, x, y, st); break;
        case EW_AXPBY: k_axpby(n, make_double2(w.a, 0.0), x, make_double2(w.b, 0.0), y, st); break;
        default: k_fill_zero(y, n, st); break;
    }
    return 0;
}

int run_denom(BlkEngine* E, int i, hipStream_t st, std::string& err) {
    const DenomOp& d = E->P.denoms[i];
    if (!E->eps_ok) {
        err = "blocked engine: orbital energies not set";
        return -1;
    }
    const int mask = (d.mask && E->P.any_frozen) ? 1 : 0;
    k_bdenom(E->base[d.S1], E->D1, E->F1m, E->P.arr[A_D1].size, d.divide ? 1 : 0, mask, st);
    k_bdenom(E->base[d.S2], E->D2, E->F2m, E->P.arr[A_D2].size, d.divide ? 1 : 0, mask, st);
    return 0;
}

int run_prog(BlkEngine* E, int prog, int gauss, hipStream_t st, std::string& err) {
    for (const Op& op : E->P.prog[prog]) {
        int rc = 0;
        switch (op.kind) {
            case OK_PROD: rc = run_prod(E, op.idx, gauss, st, err); break;
            case OK_GATHER: rc = run_gather(E, op.idx, st); break;
            case OK_BIL: rc = run_bil(E, op.idx, st); break;
            case OK_EW: rc = run_ew(E, op.idx, st); break;
            default: rc = run_denom(E, op.idx, st, err); break;
        }
        if (rc != 0) return -1;
    }
    return 0;
}

}  // namespace

int blk_create(const SymIn& s, BlkEngine** out, std::string& err) {
    *out = nullptr;
    auto* E = new BlkEngine();
    if (build_plan(s, E->P) != 0) {
        err = E->P.err;
        delete E;
        return 1;   // not applicable (e.g. irreps that are not self-inverse)
    }
    const Plan& P = E->P;
    E->base.assign(A_COUNT, nullptr);
    E->owned.assign(A_COUNT, false);
    for (int a = 0; a < A_COUNT; ++a) {
        if (a == A_T1 || a == A_T2 || a == A_S1 || a == A_S2 || a == A_VVVV || a == A_VOVV ||
            a == A_D1 || a == A_D2 || a == A_F1m || a == A_F2m)
            continue;
        const int64_t n = a == A_WS ? P.ws_elems : P.arr[a].size;
        if (balloc(E, &E->base[a], n * 16, err) != 0) {
            blk_destroy(E);
            return -1;
        }
        E->owned[a] = true;
    }
    E->base[A_T1] = E->base[A_T];
    E->base[A_T2] = E->base[A_T] + P.arr[A_T1].size;
    E->base[A_S1] = E->base[A_S];
    E->base[A_S2] = E->base[A_S] + P.arr[A_S1].size;
    if (balloc(E, &E->D1, P.arr[A_D1].size * 8, err) != 0 ||
        balloc(E, &E->D2, P.arr[A_D2].size * 8, err) != 0 ||
        balloc(E, &E->F1m, P.arr[A_D1].size * 8, err) != 0 ||
        balloc(E, &E->F2m, P.arr[A_D2].size * 8, err) != 0) {
        blk_destroy(E);
        return -1;
    }
    // index maps (the host copies are released once on the device)
    E->gmap.assign(P.gathers.size(), nullptr);
    for (size_t i = 0; i < P.gathers.size(); ++i) {
        GatherOp& G = E->P.gathers[i];
        if (upload_vec(E, &E->gmap[i], G.map, err) != 0) {
            blk_destroy(E);
            return -1;
        }
        std::vector<int32_t>().swap(G.map);
    }
    E->bdev.assign(P.bils.size(), BlkEngine::BilDev());
    for (size_t i = 0; i < P.bils.size(); ++i) {
        BilOp& B = E->P.bils[i];
        BlkEngine::BilDev& d = E->bdev[i];
        int rc = upload_vec(E, &d.m0, B.m0, err);
        for (int t = 0; t < B.np && rc == 0; ++t) {
            rc = upload_vec(E, &d.p[t], B.p[t], err);
            if (rc == 0) rc = upload_vec(E, &d.q[t], B.q[t], err);
        }
        if (rc != 0) {
            blk_destroy(E);
            return -1;
        }
        std::vector<int32_t>().swap(B.m0);
        for (int t = 0; t < 2; ++t) {
            std::vector<int32_t>().swap(B.p[t]);
            std::vector<int32_t>().swap(B.q[t]);
        }
    }
    *out = E;
    return 0;
}

void blk_destroy(BlkEngine* E) {
    if (!E) return;
    for (void* p : E->allocs) (void)hipFree(p);
    delete E;
}

int blk_bind(BlkEngine* E, int id, double2* p, std::string& err) {
    E->base[id] = p;
    // the products' descriptors hold array addresses: built once both
    // integral blocks that live outside the engine are bound
    if (E->base[A_VVVV] && E->base[A_VOVV] && E->pdev.empty()) return build_products(E, err);
    return 0;
}

double2* blk_ptr(BlkEngine* E, int id) { return E->base[id]; }

int64_t blk_size(const BlkEngine* E, int id) { return E->P.arr[id].size; }

int64_t blk_bytes(const BlkEngine* E) { return E->bytes; }

int blk_set_eps(BlkEngine* E, const double* eps, std::string& err) {
    std::vector<double> d1, d2, f1, f2;
    build_denominators(E->P, eps, d1, d2, f1, f2);
    if (hipMemcpy(E->D1, d1.data(), d1.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(E->D2, d2.data(), d2.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(E->F1m, f1.data(), f1.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(E->F2m, f2.data(), f2.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
        err = "blocked engine: denominator upload failed";
        return -1;
    }
    E->eps_ok = true;
    return 0;
}

void blk_invalidate(BlkEngine* E) {
    E->tau_ok = E->ring_ok = E->x_ok = false;
}

int blk_run(BlkEngine* E, int prog, int gauss, hipStream_t st, std::string& err) {
    if (E->pdev.empty()) {
        err = "blocked engine: integral blocks not bound";
        return -1;
    }
    auto ensure = [&](int p, bool& ok) -> int {
        if (ok) return 0;
        if (p == P_X && !E->tau_ok) {
            if (run_prog(E, P_TAU, gauss, st, err) != 0) return -1;
            E->tau_ok = true;
        }
        if (run_prog(E, p, gauss, st, err) != 0) return -1;
        ok = true;
        return 0;
    };
    switch (prog) {
        case P_HOINTM: case P_HVINTM: case P_AINTM:
            if (ensure(P_TAU, E->tau_ok) != 0) return -1;
            break;
        case P_T1EQNS:
            if (ensure(P_TAU, E->tau_ok) != 0 || ensure(P_X, E->x_ok) != 0 ||
                ensure(P_RING, E->ring_ok) != 0)
                return -1;
            break;
        case P_T2EQNS:
            if (ensure(P_TAU, E->tau_ok) != 0 || ensure(P_RING, E->ring_ok) != 0) return -1;
            break;
        case P_BINTM:
            if (ensure(P_TAU, E->tau_ok) != 0 || ensure(P_X, E->x_ok) != 0) return -1;
            break;
        default: break;
    }
    if (run_prog(E, prog, gauss, st, err) != 0) return -1;
    if (prog == P_TAUS) E->tau_ok = false;     // TAU holds tau(S)
    if (prog == P_MP2) blk_invalidate(E);
    return 0;
}

void blk_tau_s_params(const BlkEngine* E, void* q) {
    for (const Op& op : E->P.prog[P_TAUS])
        if (op.kind == OK_BIL) bil_args(E, op.idx, q);
}

double blk_iteration_flops(const BlkEngine* E) {
    double f = 0.0;
    for (int p : {P_X, P_HOINTM, P_HVINTM, P_HOVINTM, P_T1EQNS, P_GOINTM, P_GVINTM, P_AINTM,
                  P_HINTM, P_T2EQNS, P_BINTM})
        for (const Op& op : E->P.prog[p])
            if (op.kind == OK_PROD) f += E->P.prods[op.idx].flops;
    return f;
}

}  // namespace rcc
