// Integral front end: MOLFDIR MRCONEE + MDCINT -> the six antisymmetrised
// classes and the Fock blocks, on the GPU.  Replaces RDINTS (ccints.F:33-569)
// with MKINDEX (:1421), GET2INTS (:765) and GETFM / MAKEFM (:1031-1419) for
// the single-reference (non Fock-space) case.
//
// The host parses the two Fortran unformatted files (4-byte record markers).
// The Kramers-reduced list is uploaded as one array of canonical integrals
// (IKR JKR | KKR LKR) and expanded on the device, one thread per canonical
// integral, into its 16 index quadruples in GET2INTS's loop order: time
// reversal of the (IJ) and (KL) pairs, particle interchange and complex
// conjugation.  Each (ij|kl) = <ik|jl> belongs to the class of the (ik; jl)
// pair types -- the direct part when (i, k) is a stored pair, the exchange
// part at (k, i) otherwise.
//
// relccsd stores every expanded value by assignment while it reads the file,
// so when two expansions land on one element the later one in file order
// is kept.  Two expansions land on one (element, part) only if they are the
// same index quadruple, and the 16 expansions of a list entry are its whole
// orbit under the pair time reversals, particle interchange and conjugation
// -- the same set of quadruples for every entry of that orbit.  So the last
// entry of each orbit in file order writes every element its orbit reaches,
// and within that entry the last expansion wins.  The device finds those
// entries without ordering the images: a claim pass takes, per orbit, the
// largest entry index (one 32-bit atomicMax per entry).  The orbit's key is
// its canonical member's index: every orbit of Kramers tuples (a, b | c, d)
// under the 16 operations holds tuples with a > 0, |b| <= a, 0 < |c| <= a,
// |d| <= |c|, and the smallest enumeration index among them names the orbit.
// The keys of one a (the largest Kramers index of the orbit, constant on
// it) fill 4 a^2 (a + 1) consecutive words, so a list written in IKR order
// walks them through the caches.  Where that array would exceed
// SORT_CANON_MAX bytes the key is the orbit's anchor instead -- the smallest
// (class, part, element) among its stored expansions -- kept in a word of
// the still-zero class element, and a flag pass marks the holders and
// clears the keys before any class write.  Only the holders write, each
// element and part then from one thread, in two ordered passes of plain
// stores: the direct pass stores d into the zeroed class, and the exchange
// pass, launched after it, replaces the element by fl(d - e) -- RDINTS's
// sequential stores followed by its direct - exchange XAXPY.
#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <string>
#include <vector>
#include "rcc_internal.h"

namespace rcc {

namespace {

struct Reader {
    FILE* f = nullptr;
    std::string path;
    ~Reader() {
        if (f) fclose(f);
    }
    // next record into buf; false at end of file
    bool next(std::vector<unsigned char>& buf, std::string& err) {
        int32_t n = 0, m = 0;
        if (fread(&n, 4, 1, f) != 1) return false;
        if (n < 0) {
            err = path + ": records longer than 2 GiB (split records) are not supported";
            return false;
        }
        buf.resize((size_t)n);
        if (n && fread(buf.data(), 1, (size_t)n, f) != (size_t)n) {
            err = path + ": truncated record";
            return false;
        }
        if (fread(&m, 4, 1, f) != 1 || m != n) {
            err = path + ": bad record marker";
            return false;
        }
        return true;
    }
};

template <class T>
T get(const std::vector<unsigned char>& b, size_t& p) {
    T v;
    std::memcpy(&v, b.data() + p, sizeof(T));
    p += sizeof(T);
    return v;
}

}  // namespace

int read_mrconee(const char* path, MolfdirOne& out, std::string& err) {
    Reader r;
    r.path = path;
    r.f = fopen(path, "rb");
    if (!r.f) {
        err = std::string("cannot open ") + path;
        return -1;
    }
    std::vector<unsigned char> b;
    size_t p = 0;
    // 1: NSP, BREIT, ECORE
    if (!r.next(b, err) || b.size() < 16) {
        if (err.empty()) err = std::string(path) + ": record 1";
        return -1;
    }
    out.nsp = get<int32_t>(b, p);
    out.breit = get<int32_t>(b, p) != 0;
    out.ecore = get<double>(b, p);
    // 2: NSYMRP + names, 3: NREP + abelian names, 4: MULTB
    if (!r.next(b, err) || b.size() < 4) {
        if (err.empty()) err = std::string(path) + ": record 2";
        return -1;
    }
    p = 0;
    out.nsymrp = get<int32_t>(b, p);
    if (!r.next(b, err) || b.size() < 4) {
        if (err.empty()) err = std::string(path) + ": record 3";
        return -1;
    }
    p = 0;
    out.nrep = get<int32_t>(b, p);
    if (!r.next(b, err)) return -1;
    // 5: per spinor (higher-group irrep, abelian irrep, orbital energy)
    if (!r.next(b, err) || b.size() != (size_t)out.nsp * 16) {
        if (err.empty()) err = std::string(path) + ": record 5 size";
        return -1;
    }
    p = 0;
    out.irrep_higher.resize(out.nsp);
    out.irrep.resize(out.nsp);
    out.eps.resize(out.nsp);
    for (int s = 0; s < out.nsp; ++s) {
        out.irrep_higher[s] = get<int32_t>(b, p);
        out.irrep[s] = get<int32_t>(b, p);
        out.eps[s] = get<double>(b, p);
    }
    // 6: the one-electron matrix, (Re, Im), column-major
    if (!r.next(b, err) || b.size() != (size_t)out.nsp * out.nsp * 16) {
        if (err.empty()) err = std::string(path) + ": record 6 (one-electron matrix) missing or of wrong size";
        return -1;
    }
    out.h1.resize((size_t)out.nsp * out.nsp);
    std::memcpy(out.h1.data(), b.data(), b.size());
    return 0;
}

int read_mdcint(const char* path, MdcintList& out, std::string& err, bool header_only_ok) {
    Reader r;
    r.path = path;
    r.f = fopen(path, "rb");
    if (!r.f) {
        err = std::string("cannot open ") + path;
        return -1;
    }
    std::vector<unsigned char> b;
    if (!r.next(b, err) || b.size() < 22) {
        if (err.empty()) err = std::string(path) + ": header record";
        return -1;
    }
    size_t p = 18;     // DATEX (10) + TIMEX (8)
    out.nkr = get<int32_t>(b, p);
    const bool bare = b.size() == 22;     // MDBINT: RDINTS reads DATEX, TIMEX, NKR only
    if (out.nkr <= 0 || (b.size() != 22 + (size_t)out.nkr * 8 && !(bare && header_only_ok))) {
        err = std::string(path) + ": header record size does not match NKR";
        return -1;
    }
    out.kr.assign(2 * (size_t)out.nkr, 0);
    if (!bare) std::memcpy(out.kr.data(), b.data() + 22, out.kr.size() * 4);
    out.idx.clear();
    out.val.clear();
    int continuations = 0;
    while (true) {
        if (!r.next(b, err)) {
            if (err.empty()) err = r.path + ": no end record (IKR = JKR = 0)";
            return -1;
        }
        if (b.size() < 12) {
            err = r.path + ": short integral record";
            return -1;
        }
        p = 0;
        const int ikr = get<int32_t>(b, p), jkr = get<int32_t>(b, p), nz = get<int32_t>(b, p);
        if (ikr == 0) {
            if (jkr == 0) break;
            // continuation (GET2INTS ccints.F:832-838): the next record names
            // the file (CHARACTER*8) whose integral records follow, without a
            // header; it is looked for beside the file that names it
            if (!r.next(b, err) || b.size() != 8) {
                if (err.empty()) err = r.path + ": continuation record without a file name";
                return -1;
            }
            std::string next((const char*)b.data(), 8);
            next.erase(next.find_last_not_of(' ') + 1);
            const std::string cur = r.path;
            const size_t slash = cur.find_last_of('/');
            const std::string full =
                (next.empty() || next[0] == '/' || slash == std::string::npos)
                    ? next : cur.substr(0, slash + 1) + next;
            if (++continuations > 4096 || next.empty()) {
                err = cur + ": bad continuation file name or too many continuation files";
                return -1;
            }
            fclose(r.f);
            r.path = full;
            r.f = fopen(full.c_str(), "rb");
            if (!r.f) {
                err = "cannot open continuation file " + full + " (named in " + cur + ")";
                return -1;
            }
            continue;
        }
        if (nz < 0 || b.size() != 12 + (size_t)nz * 24) {
            err = r.path + ": integral record size does not match NZ";
            return -1;
        }
        const size_t base = out.idx.size();
        out.idx.resize(base + nz);
        out.val.resize(base + nz);
        for (int z = 0; z < nz; ++z) {
            int32_t kl[2];
            std::memcpy(kl, b.data() + 12 + (size_t)z * 8, 8);
            out.idx[base + z] = make_int4(ikr, jkr, kl[0], kl[1]);
        }
        std::memcpy(out.val.data() + base, b.data() + 12 + (size_t)nz * 8, (size_t)nz * 16);
    }
    return 0;
}

int read_mdprop(const char* path, const char* name, int nsp, std::vector<double2>& prop,
                std::string& err) {
    Reader r;
    r.path = path;
    r.f = fopen(path, "rb");
    if (!r.f) {
        err = std::string("cannot open ") + path;
        return -1;
    }
    std::string want(name);
    want.resize(8, ' ');
    std::vector<unsigned char> b;
    while (r.next(b, err)) {
        if (b.size() != 32 || std::memcmp(b.data(), "********", 8) != 0 ||
            std::memcmp(b.data() + 24, want.data(), 8) != 0)
            continue;
        if (!r.next(b, err) || b.size() != (size_t)nsp * nsp * 16) {
            if (err.empty())
                err = std::string(path) + ": operator " + name + " is not an NSP x NSP matrix";
            return -1;
        }
        prop.resize((size_t)nsp * nsp);
        std::memcpy(prop.data(), b.data(), b.size());
        return 0;
    }
    if (err.empty()) err = std::string("property ") + name + " not found on " + path;
    return -1;
}

namespace {

struct SortDev {
    int no, nv;
    int itrsgn;                      // GET2INTS's ITRSGN: 1 Coulomb, -1 Breit
    int64_t rows[6];
    double2* cls[6];
    unsigned* ckey;                  // orbit keys by canonical index, or null (anchor keys)
    int all_last;                    // every entry is the last of its orbit (no duplicates)
};

// canonical orbit keys only up to this many bytes of key array
constexpr int64_t SORT_CANON_MAX = 4ll << 30;

// number of canonical tuples with largest index below a: sum_{x<a} 4 x^2 (x+1)
__host__ __device__ __forceinline__ int64_t canon_base(int64_t a) {
    const int64_t n = a - 1;
    const int64_t s3 = (n * (n + 1) / 2) * (n * (n + 1) / 2);
    const int64_t s2 = n * (n + 1) * (2 * n + 1) / 6;
    return 4 * (s3 + s2);
}

// enumeration index of a canonical tuple: a, then b (2a values), then c in
// the order 1, -1, 2, -2, .., then d (2|c| values), negative before positive
__device__ __forceinline__ int64_t canon_index(int a, int b, int c, int d) {
    const int ac = c < 0 ? -c : c;
    const int bi = b < 0 ? b + a : b + a - 1;
    const int ci = 2 * ac * (ac - 1) + (c > 0 ? 0 : 2 * ac);
    const int di = d < 0 ? d + ac : d + ac - 1;
    return canon_base(a) + (int64_t)bi * 2 * a * (a + 1) + ci + di;
}

__device__ __forceinline__ bool canon_form(int a, int b, int c, int d) {
    const int ab = b < 0 ? -b : b, ac = c < 0 ? -c : c, ad = d < 0 ? -d : d;
    return a > 0 && ab <= a && ac <= a && ad <= ac;
}

// the orbit key of a list entry: the smallest canonical index among its 16
// tuples, in the order of GET2INTS's loops (ccints.F:936-1021)
__device__ __forceinline__ int64_t orbit_key(int4 q) {
    int64_t best = INT64_MAX;
#pragma unroll
    for (int t1 = 0; t1 < 2; ++t1) {
        const int i0 = t1 ? -q.y : q.x, j0 = t1 ? -q.x : q.y;
#pragma unroll
        for (int t2 = 0; t2 < 2; ++t2) {
            const int k0 = t2 ? -q.w : q.z, l0 = t2 ? -q.z : q.w;
#pragma unroll
            for (int pi = 0; pi < 2; ++pi) {
                const int i1 = pi ? k0 : i0, j1 = pi ? l0 : j0;
                const int k1 = pi ? i0 : k0, l1 = pi ? j0 : l0;
#pragma unroll
                for (int cc = 0; cc < 2; ++cc) {
                    const int a = cc ? j1 : i1, b = cc ? i1 : j1;
                    const int c = cc ? l1 : k1, d = cc ? k1 : l1;
                    if (canon_form(a, b, c, d)) {
                        const int64_t x = canon_index(a, b, c, d);
                        best = x < best ? x : best;
                    }
                }
            }
        }
    }
    return best;
}

// MKINDEX for 0-based spinors s: spos[s] >= 0 occupied position, else
// virtual position -spos - 1
__device__ __forceinline__ int ptype(int pi, int pj) {   // 1 oo, 2 vo, 3 ov, 4 vv
    return (pi >= 0 ? 1 : 2) + (pj >= 0 ? 0 : 2);
}

__device__ __forceinline__ int64_t tri(int p, int q, int n) {
    return (int64_t)q * n - ((int64_t)q * (q + 1)) / 2 + (p - q - 1);
}

// packed pair index + 1 of (I, J), 0 when not a stored pair
__device__ __forceinline__ int64_t pindex(int pi, int pj, int no, int nv) {
    const int t = ptype(pi, pj);
    if (t == 1) return pi > pj ? tri(pi, pj, no) + 1 : 0;
    if (t == 4) {
        const int a = -pi - 1, b = -pj - 1;
        return a > b ? tri(a, b, nv) + 1 : 0;
    }
    if (t == 2) return (int64_t)pj * nv + (-pi - 1) + 1;
    return 0;
}

__device__ __forceinline__ int class_of(int t) {
    switch (t) {
        case 1: return 0;    // OOOO
        case 2: return 1;    // VOOO
        case 4: return 2;    // VVOO
        case 6: return 3;    // VOVO
        case 14: return 4;   // VOVV
        case 16: return 5;   // VVVV
        default: return -1;
    }
}

// The 16 expansions of one canonical integral in GET2INTS's loop order
// (ccints.F:936-1021); f(v, class, part, element, re sign, im sign).
template <class F>
__device__ __forceinline__ void expand16(int4 q, const int* __restrict__ kr, int nkr,
                                         const int* __restrict__ spos, const SortDev& d, F f) {
    const int IK = q.x, JK = q.y, KK = q.z, LK = q.w;
    const int ijsign = ((IK > 0) == (JK > 0) ? 1 : -1) * d.itrsgn;     // ccints.F:855
    int i0 = kr[nkr + IK], j0 = kr[nkr + JK];
    int ijsg = 1;
    int v = 0;
#pragma unroll
    for (int t1 = 0; t1 < 2; ++t1) {
        int k0 = kr[nkr + KK], l0 = kr[nkr + LK];
        int klsg = 1;
#pragma unroll
        for (int t2 = 0; t2 < 2; ++t2) {
            int i1 = i0, j1 = j0, k1 = k0, l1 = l0;
#pragma unroll
            for (int pi = 0; pi < 2; ++pi) {
                int i = i1, j = j1, k = k1, l = l1;
#pragma unroll
                for (int cc = 0; cc < 2; ++cc, ++v) {
                    const int sj = spos[j], sl = spos[l];
                    const int64_t jl = pindex(sj, sl, d.no, d.nv);
                    if (jl != 0 && i != k) {
                        const int si = spos[i], sk = spos[k];
                        const int64_t ik = pindex(si, sk, d.no, d.nv);
                        int part, t;
                        int64_t row;
                        if (ik != 0) {
                            part = 0;
                            t = (ptype(sj, sl) - 1) * 4 + ptype(si, sk);
                            row = ik - 1;
                        } else {
                            part = 1;
                            t = (ptype(sj, sl) - 1) * 4 + ptype(sk, si);
                            row = pindex(sk, si, d.no, d.nv) - 1;
                        }
                        const int c = class_of(t);
                        if (c >= 0) {
                            const int sg = ijsg * klsg;
                            f(v, c, part, row + (jl - 1) * d.rows[c], sg, cc == 0 ? sg : -sg);
                        }
                    }
                    i = j1;
                    j = i1;
                    k = l1;
                    l = k1;
                }
                i1 = k0;
                j1 = l0;
                k1 = i0;
                l1 = j0;
            }
            k0 = kr[nkr - LK];
            l0 = kr[nkr - KK];
            klsg = ((KK > 0) == (LK > 0) ? 1 : -1) * d.itrsgn;      // ccints.F:939
        }
        i0 = kr[nkr - JK];
        j0 = kr[nkr - IK];
        ijsg = ijsign;
    }
}

__device__ __forceinline__ bool in_range(int x, int nkr) { return x != 0 && x >= -nkr && x <= nkr; }

__device__ __forceinline__ bool valid(int4 q, int nkr) {
    return in_range(q.x, nkr) && in_range(q.y, nkr) && in_range(q.z, nkr) && in_range(q.w, nkr);
}

// An expansion as one 64-bit code: class (bits 61-63), part (bit 60: 0
// direct, 1 exchange), element (bits 0-59); NOIMG for one that is not stored.
constexpr unsigned long long NOIMG = ~0ull;

__device__ __forceinline__ unsigned long long img_code(int c, int part, int64_t e) {
    return ((unsigned long long)c << 61) | ((unsigned long long)part << 60) |
           (unsigned long long)e;
}

struct Images {
    unsigned long long code[16];
    signed char sre[16], sim[16];
};

// the 16 expansions of an entry in GET2INTS's order (registers: every index
// is a constant once the expansion loops are unrolled)
__device__ __forceinline__ void images16(int4 q, const int* __restrict__ kr, int nkr,
                                         const int* __restrict__ spos, const SortDev& d,
                                         Images& im) {
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        im.code[v] = NOIMG;
        im.sre[v] = 1;
        im.sim[v] = 1;
    }
    expand16(q, kr, nkr, spos, d, [&](int v, int c, int part, int64_t e, int sre, int sim) {
        im.code[v] = img_code(c, part, e);
        im.sre[v] = (signed char)sre;
        im.sim[v] = (signed char)sim;
    });
}

__device__ __forceinline__ unsigned long long anchor_of(const Images& im) {
    unsigned long long a = NOIMG;
#pragma unroll
    for (int v = 0; v < 16; ++v) a = im.code[v] < a ? im.code[v] : a;
    return a;
}

// the 32-bit key of an anchor: word `part` of its (zero until the writes)
// class element
__device__ __forceinline__ unsigned* anchor_key(const SortDev& d, unsigned long long a) {
    const int c = (int)(a >> 61), part = (int)((a >> 60) & 1ull);
    const int64_t e = (int64_t)(a & ((1ull << 60) - 1ull));
    return reinterpret_cast<unsigned*>(d.cls[c] + e) + part;
}

// Claim: per anchor, the largest 1-based entry index.  Entries with a
// Kramers index outside +-NKR are skipped and flagged.
__global__ void sort_claim_kernel(SortDev d, const int4* __restrict__ idx, int64_t m,
                                  const int* __restrict__ kr, int nkr,
                                  const int* __restrict__ spos, int* __restrict__ bad) {
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < m;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int4 q = idx[n];
        if (!valid(q, nkr)) {
            atomicOr(bad, 1);
            continue;
        }
        if (d.ckey) {
            const int64_t k = orbit_key(q);
            if (k == INT64_MAX) atomicOr(bad, 2);      // no canonical member (never seen)
            else atomicMax(d.ckey + k, (unsigned)(n + 1));
            continue;
        }
        Images im;
        images16(q, kr, nkr, spos, d, im);
        const unsigned long long a = anchor_of(im);
        if (a != NOIMG) atomicMax(anchor_key(d, a), (unsigned)(n + 1));
    }
}

// Canonical keys, first pass: every entry stores its index into its orbit's
// key (plain stores: with two entries of one orbit either may stay).  When
// as many keys are set as there are entries, no orbit has two entries and
// every entry writes; otherwise the claim pass (atomicMax) settles the
// last one.
__global__ void sort_mark_kernel(SortDev d, const int4* __restrict__ idx, int64_t m, int nkr,
                                 int* __restrict__ bad) {
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < m;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int4 q = idx[n];
        if (!valid(q, nkr)) {
            atomicOr(bad, 1);
            continue;
        }
        const int64_t k = orbit_key(q);
        if (k == INT64_MAX) atomicOr(bad, 2);      // no canonical member (never seen)
        else d.ckey[k] = (unsigned)(n + 1);
    }
}

// number of set keys: per-wave sums by lane shuffles, one atomic per wave
__global__ void count_set_kernel(const unsigned* __restrict__ p, int64_t n,
                                 unsigned long long* __restrict__ out) {
    unsigned long long c = 0;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
         x += (int64_t)gridDim.x * blockDim.x)
        c += p[x] != 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

// whether entry n writes: every valid entry when no orbit has two, else the
// holder of its orbit's key (canonical keys) or flagged by the flag pass
// (anchor keys)
__device__ __forceinline__ bool holds(const SortDev& d, int4 q, int64_t n, int nkr,
                                      const unsigned char* __restrict__ last) {
    if (d.all_last) return valid(q, nkr);
    if (d.ckey) {
        if (!valid(q, nkr)) return false;
        const int64_t k = orbit_key(q);
        return k != INT64_MAX && d.ckey[k] == (unsigned)(n + 1);
    }
    return last[n] != 0;
}

// Flags the last entry of every orbit (the holder of its anchor's key) and
// clears the key, so the classes are zero again; an entry that is not the
// holder reads either the holder's index or 0, neither its own.
__global__ void sort_flag_kernel(SortDev d, const int4* __restrict__ idx, int64_t m,
                                 const int* __restrict__ kr, int nkr,
                                 const int* __restrict__ spos,
                                 unsigned char* __restrict__ last) {
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < m;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int4 q = idx[n];
        unsigned char w = 0;
        if (valid(q, nkr)) {
            Images im;
            images16(q, kr, nkr, spos, d, im);
            const unsigned long long a = anchor_of(im);
            if (a != NOIMG) {
                unsigned* key = anchor_key(d, a);
                if (__hip_atomic_load(key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                    (unsigned)(n + 1)) {
                    w = 1;
                    __hip_atomic_store(key, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        last[n] = w;
    }
}

// The direct pass: every flagged entry stores its direct expansions, in
// expansion order (a later one on the same element overwrites).
__global__ void sort_direct_kernel(SortDev d, const int4* __restrict__ idx,
                                   const double2* __restrict__ val, int64_t m,
                                   const int* __restrict__ kr, int nkr,
                                   const int* __restrict__ spos,
                                   const unsigned char* __restrict__ last) {
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < m;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int4 q = idx[n];
        if (!holds(d, q, n, nkr, last)) continue;
        const double2 x = val[n];
        expand16(q, kr, nkr, spos, d, [&](int, int c, int part, int64_t e, int sre, int sim) {
            if (part == 0) d.cls[c][e] = make_double2(sre * x.x, sim * x.y);
        });
    }
}

// The exchange pass, after the direct one: every flagged entry replaces each
// element its exchange expansions reach by fl(d - e), with e the last of its
// expansions on that element.
__global__ void sort_exchange_kernel(SortDev d, const int4* __restrict__ idx,
                                     const double2* __restrict__ val, int64_t m,
                                     const int* __restrict__ kr, int nkr,
                                     const int* __restrict__ spos,
                                     const unsigned char* __restrict__ last) {
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < m;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int4 q = idx[n];
        if (!holds(d, q, n, nkr, last)) continue;
        const double2 x = val[n];
        Images im;
        images16(q, kr, nkr, spos, d, im);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const unsigned long long cv = im.code[v];
            if (cv == NOIMG || ((cv >> 60) & 1ull) == 0) continue;
            bool later = false;
#pragma unroll
            for (int w = v + 1; w < 16; ++w) later = later || im.code[w] == cv;
            if (later) continue;
            double2* p = d.cls[(int)(cv >> 61)] + (int64_t)(cv & ((1ull << 60) - 1ull));
            const double2 o = *p;
            *p = make_double2(o.x - im.sre[v] * x.x, o.y - im.sim[v] * x.y);
        }
    }
}

__global__ void fill_u32_kernel(unsigned* p, int64_t n) {
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
         x += (int64_t)gridDim.x * blockDim.x)
        p[x] = 0u;
}

__global__ void fill_z_kernel(double2* p, int64_t n) {
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
         x += (int64_t)gridDim.x * blockDim.x)
        p[x] = make_double2(0.0, 0.0);
}

int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1 << 16)); }

// GETFM + MAKEFM (ccints.F:1031-1419) on packed C1 classes:
//   F(i,j) = conj h(I,J) + sum_k <ik||jk>,  F(a,i) = conj h + sum_k <ak||ik>,
//   F(a,c) = conj h + sum_k <ak||ck>
// h1 is indexed by spinor; occ / vir hold the spinor of each position.
__global__ void fock_kernel(int no, int nv, int nsp, const double2* __restrict__ h1,
                            const int* __restrict__ occ, const int* __restrict__ vir,
                            const double2* __restrict__ OOOO, const double2* __restrict__ VOOO,
                            const double2* __restrict__ VOVO, double2* __restrict__ FOO,
                            double2* __restrict__ FVO, double2* __restrict__ FVV) {
    const int64_t noot = (int64_t)no * (no - 1) / 2, nvo = (int64_t)no * nv;
    const int64_t total = (int64_t)no * no + nvo + (int64_t)nv * nv;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total;
         x += (int64_t)gridDim.x * blockDim.x) {
        double2 f;
        if (x < (int64_t)no * no) {
            const int i = (int)(x % no), j = (int)(x / no);
            f = h1[occ[i] + (int64_t)occ[j] * nsp];
            f.y = -f.y;
            for (int k = 0; k < no; ++k) {
                if (k == i || k == j) continue;
                // <ik||jk>: bra pair (i,k), ket pair (j,k), one sign per reversed pair
                const int64_t b = i > k ? tri(i, k, no) : tri(k, i, no);
                const int64_t c = j > k ? tri(j, k, no) : tri(k, j, no);
                const double s = (i > k) == (j > k) ? 1.0 : -1.0;
                const double2 v = OOOO[b + c * noot];
                f.x += s * v.x;
                f.y += s * v.y;
            }
            FOO[x] = f;
        } else if (x < (int64_t)no * no + nvo) {
            const int64_t y = x - (int64_t)no * no;
            const int a = (int)(y % nv), i = (int)(y / nv);
            f = h1[vir[a] + (int64_t)occ[i] * nsp];
            f.y = -f.y;
            for (int k = 0; k < no; ++k) {
                if (k == i) continue;
                const int64_t c = i > k ? tri(i, k, no) : tri(k, i, no);
                const double s = i > k ? 1.0 : -1.0;
                const double2 v = VOOO[(int64_t)k * nv + a + c * nvo];
                f.x += s * v.x;
                f.y += s * v.y;
            }
            FVO[y] = f;
        } else {
            const int64_t y = x - (int64_t)no * no - nvo;
            const int a = (int)(y % nv), c = (int)(y / nv);
            f = h1[vir[a] + (int64_t)vir[c] * nsp];
            f.y = -f.y;
            for (int k = 0; k < no; ++k) {
                const double2 v = VOVO[(int64_t)k * nv + a + ((int64_t)k * nv + c) * nvo];
                f.x += v.x;
                f.y += v.y;
            }
            FVV[y] = f;
        }
    }
}

// E0 = sum_i Re F(i,i), E1 = -1/2 sum_ik Re <ik||ik>, eps = Re diag F (one
// workgroup, fixed order)
__global__ void __launch_bounds__(256)
fock_energy_kernel(int no, int nv, const double2* __restrict__ OOOO,
                   const double2* __restrict__ FOO, const double2* __restrict__ FVV,
                   double* __restrict__ eps, double* __restrict__ e01) {
    __shared__ double sh[2][4];
    const int tid = threadIdx.x;
    const int64_t noot = (int64_t)no * (no - 1) / 2;
    double e0 = 0.0, e1 = 0.0;
    for (int i = tid; i < no; i += 256) {
        e0 += FOO[i + (int64_t)i * no].x;
        eps[i] = FOO[i + (int64_t)i * no].x;
    }
    for (int a = tid; a < nv; a += 256) eps[no + a] = FVV[a + (int64_t)a * nv].x;
    for (int64_t p = tid; p < noot; p += 256) e1 -= OOOO[p + p * noot].x;
    // wave sums by shuffles (fixed order), then the four waves in order
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        e0 += __shfl_down(e0, off, 64);
        e1 += __shfl_down(e1, off, 64);
    }
    if ((tid & 63) == 0) {
        sh[0][tid >> 6] = e0;
        sh[1][tid >> 6] = e1;
    }
    __syncthreads();
    if (tid == 0) {
        e01[0] = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
        e01[1] = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
    }
}

}  // namespace

#define FECK(call)                                                              \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) {                                                 \
            err = std::string(#call) + ": " + hipGetErrorString(e_);            \
            return -1;                                                          \
        }                                                                       \
    } while (0)

int sort_integrals_device(int no, int nv, double2* const* cls, const int* d_spos, int nsp,
                          const int* d_kr, int nkr, const int4* d_idx, const double2* d_val,
                          int64_t m, hipStream_t st, double* ms, std::string& err, int itrsgn) {
    (void)nsp;
    SortDev d;
    d.no = no;
    d.nv = nv;
    d.itrsgn = itrsgn;
    const int64_t noot = (int64_t)no * (no - 1) / 2, nvvt = (int64_t)nv * (nv - 1) / 2,
                  nvo = (int64_t)no * nv;
    const int64_t rows[6] = {noot, nvo, nvvt, nvo, nvo, nvvt};
    const int64_t cols[6] = {noot, noot, noot, nvo, nvvt, nvvt};
    for (int c = 0; c < 6; ++c) {
        d.rows[c] = rows[c];
        d.cls[c] = cls[c];
    }
    if (m >= 0xffffffffll) {
        err = "integral sort: more than 2^32 - 2 canonical integrals";
        return -1;
    }
    // canonical orbit keys when the array is small enough (and allocatable);
    // RCC_SORT_KEYS=anchor in the environment takes the anchor keys always
    const int64_t nckey = canon_base((int64_t)nkr + 1);
    const char* ek = std::getenv("RCC_SORT_KEYS");
    const bool anchor_only = ek && std::strcmp(ek, "anchor") == 0;
    d.ckey = nullptr;
    if (!anchor_only && nckey * 4 <= SORT_CANON_MAX) {
        if (hipMalloc((void**)&d.ckey, (size_t)std::max<int64_t>(nckey, 1) * 4) != hipSuccess) {
            (void)hipGetLastError();
            d.ckey = nullptr;
        }
    }
    d.all_last = 0;
    unsigned char* last = nullptr;
    // bad flags (int) and the set-key count (u64) in one small buffer
    struct SortStat { int bad; int pad; unsigned long long set; };
    SortStat* stat = nullptr;
    if (!d.ckey) FECK(hipMalloc((void**)&last, (size_t)std::max<int64_t>(m, 1)));
    FECK(hipMalloc((void**)&stat, sizeof(SortStat)));
    FECK(hipMemsetAsync(stat, 0, sizeof(SortStat), st));
    int* bad = &stat->bad;
    hipEvent_t e0, e1;
    FECK(hipEventCreate(&e0));
    FECK(hipEventCreate(&e1));
    FECK(hipEventRecord(e0, st));
    for (int c = 0; c < 6; ++c)
        fill_z_kernel<<<grid_for(rows[c] * cols[c]), 256, 0, st>>>(cls[c], rows[c] * cols[c]);
    if (d.ckey) fill_u32_kernel<<<grid_for(nckey), 256, 0, st>>>(d.ckey, nckey);
    const int g = grid_for(m);
    SortStat hs{0, 0, 0};
    if (m > 0 && d.ckey) {
        sort_mark_kernel<<<g, 256, 0, st>>>(d, d_idx, m, nkr, bad);
        count_set_kernel<<<1024, 256, 0, st>>>(d.ckey, nckey, &stat->set);
        FECK(hipMemcpyAsync(&hs, stat, sizeof(SortStat), hipMemcpyDeviceToHost, st));
        FECK(hipStreamSynchronize(st));
        d.all_last = hs.bad == 0 && hs.set == (unsigned long long)m;
    }
    if (m > 0 && !hs.bad) {
        if (!d.all_last) sort_claim_kernel<<<g, 256, 0, st>>>(d, d_idx, m, d_kr, nkr, d_spos, bad);
        if (!d.ckey) sort_flag_kernel<<<g, 256, 0, st>>>(d, d_idx, m, d_kr, nkr, d_spos, last);
        sort_direct_kernel<<<g, 256, 0, st>>>(d, d_idx, d_val, m, d_kr, nkr, d_spos, last);
        sort_exchange_kernel<<<g, 256, 0, st>>>(d, d_idx, d_val, m, d_kr, nkr, d_spos, last);
    }
    FECK(hipGetLastError());
    FECK(hipEventRecord(e1, st));
    FECK(hipEventSynchronize(e1));
    float t = 0.f;
    FECK(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    int hbad = 0;
    FECK(hipMemcpy(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (last) FECK(hipFree(last));
    if (d.ckey) FECK(hipFree(d.ckey));
    FECK(hipFree(stat));
    if (hbad & 1) {
        err = "integral sort: a Kramers index of the list is 0 or outside +-NKR (entries skipped)";
        return -1;
    }
    if (hbad) {
        err = "integral sort: an entry without a canonical form (entries skipped)";
        return -1;
    }
    return 0;
}

int build_fock_device(int no, int nv, int nsp, const double2* d_h1, const int* d_occ,
                      const int* d_vir, double2* const* cls, double2* FOO, double2* FVO,
                      double2* FVV, double* eps, double* d_e01, hipStream_t st, std::string& err) {
    const int64_t total = (int64_t)no * no + (int64_t)no * nv + (int64_t)nv * nv;
    fock_kernel<<<grid_for(total), 256, 0, st>>>(no, nv, nsp, d_h1, d_occ, d_vir, cls[0], cls[1],
                                                 cls[3], FOO, FVO, FVV);
    FECK(hipGetLastError());
    fock_energy_kernel<<<1, 256, 0, st>>>(no, nv, cls[0], FOO, FVV, eps, d_e01);
    FECK(hipGetLastError());
    return 0;
}

}  // namespace rcc

// ---- MCCRES restart file (PUTAMPT / GETAMPT ccgetv.F:577-709, CCDUMP
// ccdriv.F:2767-2816, CCDINI ccdriv.F:2744-2751) ----
// Direct access, records of RSRCLEN = 1024 REAL*8 (8192 bytes; IRECLEN for
// LINUX, gpluuk.F:99-128): record 1 input (unused, zero), 2 IMSTAT(30),
// 3 ETOT ESCF EMP2 ECCSD ET1 ET2 ET3, from 4 the T1 amplitudes and from
// 4 + (2 NDIMT1 / 1024 + 1) the T2 amplitudes, 512 complex per record.
namespace {

thread_local std::string g_file_err;
constexpr int64_t RSRCLEN = 1024;

int64_t nrec(int64_t ndim) { return ndim * 2 / RSRCLEN + 1; }

}  // namespace

extern "C" const char* rcc_file_error(void) { return g_file_err.c_str(); }

// MRCONEE records 1-6 as RDSYMI (ccdriv.F:2978-2988) and GETFM (ccints.F:1091-1096)
// read them; the higher point group is written equal to the abelian one
// (names A01.. / a01..), as for the files the reference test inputs use.
extern "C" int rcc_mrconee_write(const char* path, int nsp, int nrep, const int* irreps,
                                 const int* nelec, const int* multb, const double* eps,
                                 const double* h1, double ecore, int breit) {
    if (!path || nsp < 1 || nrep < 1 || nrep > 16 || !irreps || !nelec || !eps || !h1 ||
        (!multb && nrep != 1)) {
        g_file_err = "rcc_mrconee_write: bad arguments (multb may be null only for nrep = 1)";
        return -1;
    }
    for (int s = 0; s < nsp; ++s)
        if (irreps[s] < 1 || irreps[s] > nrep) {
            g_file_err = "rcc_mrconee_write: spinor irrep out of 1..nrep";
            return -1;
        }
    FILE* f = fopen(path, "wb");
    if (!f) {
        g_file_err = std::string("cannot create ") + path;
        return -1;
    }
    bool ok = true;
    auto record = [&](const std::vector<unsigned char>& b) {
        const int32_t n = (int32_t)b.size();
        ok = ok && fwrite(&n, 4, 1, f) == 1 && (b.empty() || fwrite(b.data(), 1, b.size(), f) == b.size()) &&
             fwrite(&n, 4, 1, f) == 1;
    };
    auto put = [](std::vector<unsigned char>& b, const void* p, size_t n) {
        const unsigned char* c = (const unsigned char*)p;
        b.insert(b.end(), c, c + n);
    };
    std::vector<unsigned char> b;
    const int32_t nsp32 = nsp, brt = breit ? 1 : 0, nrep32 = nrep;
    put(b, &nsp32, 4);
    put(b, &brt, 4);
    put(b, &ecore, 8);
    record(b);
    b.clear();
    put(b, &nrep32, 4);
    char name[32];
    for (int i = 0; i < nrep; ++i) {
        std::string nm(14, ' ');                                      // character*14
        snprintf(name, sizeof name, "A%02d", i + 1);
        nm.replace(0, std::strlen(name), name);
        put(b, nm.data(), 14);
    }
    for (int i = 0; i < nrep; ++i) {
        const int32_t ne = nelec[i];
        put(b, &ne, 4);
    }
    record(b);
    b.clear();
    put(b, &nrep32, 4);
    for (int half = 0; half < 2; ++half)
        for (int i = 0; i < nrep; ++i) {
            snprintf(name, sizeof name, " %c%02d", half ? 'a' : 'A', i + 1);  // character*4
            put(b, name, 4);
        }
    record(b);
    b.clear();
    for (int j = 0; j < 2 * nrep; ++j)
        for (int i = 0; i < 2 * nrep; ++i) {
            const int32_t m = multb ? multb[i + 2 * nrep * j] : (i == j ? 2 : 1);
            put(b, &m, 4);
        }
    record(b);
    b.clear();
    for (int sp = 0; sp < nsp; ++sp) {
        const int32_t g = irreps[sp];
        put(b, &g, 4);
        put(b, &g, 4);
        put(b, &eps[sp], 8);
    }
    record(b);
    b.clear();
    put(b, h1, (size_t)nsp * nsp * 16);
    record(b);
    if (fclose(f) != 0) ok = false;
    if (!ok) {
        g_file_err = std::string("write error on ") + path;
        return -1;
    }
    return 0;
}

extern "C" int rcc_mccres_write(const char* path, int64_t ndimt1, int64_t ndimt2, const double* t1,
                                const double* t2, const int* imstat, const double* energies) {
    FILE* f = fopen(path, "wb");
    if (!f) {
        g_file_err = std::string("cannot create ") + path;
        return -1;
    }
    std::vector<double> rec(RSRCLEN, 0.0);
    bool ok = fwrite(rec.data(), 8, RSRCLEN, f) == (size_t)RSRCLEN;          // 1
    std::vector<int32_t> st(2 * RSRCLEN, 0);
    for (int i = 0; i < 30 && imstat; ++i) st[i] = imstat[i];
    ok = ok && fwrite(st.data(), 4, 2 * RSRCLEN, f) == (size_t)(2 * RSRCLEN);  // 2
    for (int i = 0; i < 7; ++i) rec[i] = energies ? energies[i] : 0.0;
    ok = ok && fwrite(rec.data(), 8, RSRCLEN, f) == (size_t)RSRCLEN;          // 3
    for (int part = 0; part < 2 && ok; ++part) {
        const int64_t n = part == 0 ? ndimt1 : ndimt2;
        const double* src = part == 0 ? t1 : t2;
        for (int64_t r = 0; r < nrec(n) && ok; ++r) {
            std::fill(rec.begin(), rec.end(), 0.0);
            const int64_t c0 = r * (RSRCLEN / 2), nc = std::max<int64_t>(0, std::min<int64_t>(n - c0, RSRCLEN / 2));
            if (nc > 0) std::memcpy(rec.data(), src + 2 * c0, (size_t)nc * 16);
            ok = fwrite(rec.data(), 8, RSRCLEN, f) == (size_t)RSRCLEN;
        }
    }
    if (fclose(f) != 0) ok = false;
    if (!ok) {
        g_file_err = std::string("write error on ") + path;
        return -1;
    }
    return 0;
}

extern "C" int rcc_mccres_read(const char* path, int64_t ndimt1, int64_t ndimt2, double* t1,
                               double* t2, int* imstat, double* energies) {
    FILE* f = fopen(path, "rb");
    if (!f) {
        g_file_err = std::string("cannot open ") + path;
        return -1;
    }
    fseek(f, 0, SEEK_END);
    const long long size = ftell(f);
    const long long need = (3 + nrec(ndimt1) + nrec(ndimt2)) * RSRCLEN * 8;
    if (size < need) {
        fclose(f);
        g_file_err = std::string(path) + ": " + std::to_string(size) + " bytes, the amplitudes of " +
                     "this layout need " + std::to_string(need);
        return -1;
    }
    std::vector<double> rec(RSRCLEN);
    bool ok = fseek(f, RSRCLEN * 8, SEEK_SET) == 0;
    std::vector<int32_t> st(2 * RSRCLEN);
    ok = ok && fread(st.data(), 4, 2 * RSRCLEN, f) == (size_t)(2 * RSRCLEN);
    if (imstat)
        for (int i = 0; i < 30; ++i) imstat[i] = st[i];
    ok = ok && fread(rec.data(), 8, RSRCLEN, f) == (size_t)RSRCLEN;
    if (energies)
        for (int i = 0; i < 7; ++i) energies[i] = rec[i];
    for (int part = 0; part < 2 && ok; ++part) {
        const int64_t n = part == 0 ? ndimt1 : ndimt2;
        double* dst = part == 0 ? t1 : t2;
        const int64_t first = part == 0 ? 3 : 3 + nrec(ndimt1);       // 0-based record
        ok = fseek(f, (long)(first * RSRCLEN * 8), SEEK_SET) == 0;
        for (int64_t r = 0; r < nrec(n) && ok; ++r) {
            ok = fread(rec.data(), 8, RSRCLEN, f) == (size_t)RSRCLEN;
            const int64_t c0 = r * (RSRCLEN / 2), nc = std::max<int64_t>(0, std::min<int64_t>(n - c0, RSRCLEN / 2));
            if (ok && nc > 0 && dst) std::memcpy(dst + 2 * c0, rec.data(), (size_t)nc * 16);
        }
    }
    fclose(f);
    if (!ok) {
        g_file_err = std::string("read error on ") + path;
        return -1;
    }
    return 0;
}
