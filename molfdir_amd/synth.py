"""Seeded synthetic relativistic CCSD inputs in relccsd's packed layouts.

Recipe (one per system, reproducible from ``seed``):

* orbital energies: occupied ~ U[-2.0, -0.3] Eh, virtual ~ U[0.2, 5.0] Eh,
  each set sorted ascending;
* a complex rank-``rank`` factor B[P, p, r] with Re and Im ~ N(0, scale^2),
  made Hermitian in (p, r) (upper triangle drawn, diagonal real);
* Mulliken integrals (pr|qs) = sum_P B[P,p,r] B[P,q,s], physicist's
  <pq|rs> = (pr|qs), antisymmetrised <pq||rs> = <pq|rs> - <pq|sr>;
* Fock matrix f = diag(eps) (+ an optional Hermitian perturbation, in which
  case eps is re-taken as Re diag(f), as relccsd's RDINTS does when it
  rebuilds the Fock diagonal, ccints.F:1252-1287).

The six integral classes are laid out exactly as relccsd's sorted files
hold them (SYMTAB offsets ccdriv.F:1150-1187; readers ccgetv.F:89-402), for
the C1 double group (one fermion irrep):

* packed occupied pair IJ (I>J), enumerated J outer / I inner
  (DENOM cceqns.F:2408-2448); same for virtual pairs AB (A>B);
* VO pair (A,K) at K*nv + A (A fastest, IIVO ccdriv.F:1050-1069);
* every class is column-major: rows = bra pair, columns = ket pair:
  OOOO <KL||IJ>, VOOO <AK||IJ>, VVOO <AB||IJ>, VOVO <AK||CI>,
  VOVV <AK||CD>, VVVV <AB||CD>;
* Fock blocks FOO(K,I) = f_KI, FVO(A,I) = f_AI, FVV(A,C) = f_AC,
  column-major (FMTOFILE ccgetv.F:34-87).

All arrays are returned as flat complex128 tensors (interleaved re/im in
memory, i.e. the REAL*8 pairs relccsd uses with RCW=2).  The generator runs
on whatever torch device is given; large systems are built on the GPU in
column blocks so the host never holds more than one class at a time.
"""
from __future__ import annotations

import numpy as np
import torch

__all__ = ["pair_index", "draw_factors", "make_system", "class_shapes", "spinor_irreps",
           "block_maps", "to_blocks", "dirichlet_split"]


def pair_index(n: int) -> tuple[np.ndarray, np.ndarray]:
    """(p, q) with p > q for the packed pair list of an n-set, in relccsd's
    order: q outer, p inner (cceqns.F:2416-2438)."""
    p, q = [], []
    for qq in range(n):
        for pp in range(qq + 1, n):
            p.append(pp)
            q.append(qq)
    return np.asarray(p, dtype=np.int64), np.asarray(q, dtype=np.int64)


def class_shapes(no: int, nv: int) -> dict:
    noot = no * (no - 1) // 2
    nvvt = nv * (nv - 1) // 2
    nvo = nv * no
    return {
        "FOO": (no, no), "FVO": (nv, no), "FVV": (nv, nv),
        "OOOO": (noot, noot), "VOOO": (nvo, noot), "VVOO": (nvvt, noot),
        "VOVO": (nvo, nvo), "VOVV": (nvo, nvvt), "VVVV": (nvvt, nvvt),
        "T1": (nv, no), "T2": (nvvt, noot),
    }


def draw_factors(no: int, nv: int, seed: int, rank: int = 120,
                 scale: float = 0.05):
    """Orbital energies and the Hermitian factor B[P,p,r] (numpy, host)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    eps_o = np.sort(rng.uniform(-2.0, -0.3, size=no))
    eps_v = np.sort(rng.uniform(0.2, 5.0, size=nv))
    n = no + nv
    re = rng.normal(0.0, scale, size=(rank, n, n))
    im = rng.normal(0.0, scale, size=(rank, n, n))
    z = re + 1j * im
    up = np.triu(z, 1)
    b = up + np.conj(np.transpose(up, (0, 2, 1)))
    idx = np.arange(n)
    b[:, idx, idx] = re[:, idx, idx]
    return np.concatenate([eps_o, eps_v]), b


def _ket_block(bra_p, bra_q, ket_r, ket_s):
    """<pq||rs> for all (p in bra_p-set, q in bra_q-set) and a list of ket
    pairs.  bra_p: B[:, Sp, r-values] etc. are gathered by the caller:
      Bpr[P, p, k] = B[P, p, r_k], Bqs[P, q, k] = B[P, q, s_k],
      Bps[P, p, k] = B[P, p, s_k], Bqr[P, q, k] = B[P, q, r_k].
    Returns X[k, p, q]."""
    Bpr, Bqs, Bps, Bqr = bra_p[0], bra_q[0], bra_p[1], bra_q[1]
    # (pr|qs): sum_P Bpr[P,p,k] Bqs[P,q,k]  -> per ket pair k: Bpr_k^T Bqs_k
    a = Bpr.permute(2, 1, 0).contiguous()          # k, p, P
    b = Bqs.permute(2, 0, 1).contiguous()          # k, P, q
    c = Bps.permute(2, 1, 0).contiguous()
    d = Bqr.permute(2, 0, 1).contiguous()
    return torch.bmm(a, b) - torch.bmm(c, d)


def _build_class(B, prow, qrow, rcol, scol, rows_packed, device, kblock):
    """Column-major class array: rows enumerate (prow[i], qrow[i]) bra pairs,
    columns enumerate (rcol[k], scol[k]) ket pairs.  Returns flat complex128.
    prow/qrow index the full spinor set; all are 1-D long tensors."""
    nrow = prow.numel()
    ncol = rcol.numel()
    out = torch.empty((ncol, nrow), dtype=torch.complex128, device=device)
    up = torch.unique(prow)
    uq = torch.unique(qrow)
    # local positions of prow/qrow inside the unique sets
    mp = torch.searchsorted(up, prow)
    mq = torch.searchsorted(uq, qrow)
    for k0 in range(0, ncol, kblock):
        k1 = min(ncol, k0 + kblock)
        r = rcol[k0:k1]
        s = scol[k0:k1]
        Bpr = B[:, up][:, :, r]
        Bps = B[:, up][:, :, s]
        Bqs = B[:, uq][:, :, s]
        Bqr = B[:, uq][:, :, r]
        X = _ket_block((Bpr, Bps), (Bqs, Bqr), r, s)   # k, |up|, |uq|
        out[k0:k1] = X[:, mp, mq]
    return out.reshape(-1)


def spinor_irreps(sym) -> np.ndarray:
    """Fermion irrep of every spinor in global order (occupied irrep-major,
    then virtual irrep-major) for sym = (occupied per irrep, virtual per
    irrep)."""
    no_l, nv_l = sym
    return np.concatenate([np.repeat(np.arange(len(no_l)), no_l),
                           np.repeat(np.arange(len(nv_l)), nv_l)])


def make_system(no: int, nv: int, seed: int, rank: int = 120,
                scale: float = 0.05, device="cpu", fock_perturb: float = 0.0,
                classes=("OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV"),
                kblock: int | None = None, sym=None) -> dict:
    """Build one seeded system.  Returns a dict with 'eps' (float64, occupied
    then virtual), the Fock blocks and the requested integral classes as flat
    complex128 tensors on ``device`` in relccsd's C1 layout.

    sym = (occupied per irrep, virtual per irrep) gives the system the
    symmetry of the abelian double group Z2^k x {fermion, boson} (2^k
    fermion irreps, products adding the Z2^k labels): factor component P
    carries irrep P mod 2^k and B[P,p,r] is kept only where the pair irrep of
    (p, r) equals it, so <pq||rs> vanishes unless (pq) and (rs) share a pair
    irrep; a Fock perturbation is kept only inside irreps.  Spinors are then
    in global order (irrep-major within occupied and within virtual)."""
    eps, b = draw_factors(no, nv, seed, rank, scale)
    n = no + nv
    lab = None
    if sym is not None:
        no_l, nv_l = (list(map(int, x)) for x in sym)
        nrep = len(no_l)
        if len(nv_l) != nrep or nrep & (nrep - 1) or sum(no_l) != no or sum(nv_l) != nv:
            raise ValueError("sym must give 2^k irreps whose counts sum to no, nv")
        lab = spinor_irreps((no_l, nv_l))
        pair = np.bitwise_xor.outer(lab, lab)                    # pair irrep of (p, r)
        keep = pair[None, :, :] == (np.arange(rank) % nrep)[:, None, None]
        b = b * keep
    dev = torch.device(device)
    B = torch.from_numpy(b).to(dev)
    f = np.diag(eps).astype(np.complex128)
    if fock_perturb:
        rng = np.random.Generator(np.random.PCG64(seed + 7919))
        v = rng.normal(0.0, fock_perturb, (n, n)) + 1j * rng.normal(0.0, fock_perturb, (n, n))
        v = 0.5 * (v + v.conj().T)
        if lab is not None:
            v = v * (lab[:, None] == lab[None, :])
        f = f + v
        eps = np.real(np.diag(f)).copy()
    out = {"no": no, "nv": nv, "eps": torch.from_numpy(eps).to(dev)}
    if sym is not None:
        out["sym"] = (no_l, nv_l)
    o = np.arange(no)
    vv = np.arange(no, n)
    out["FOO"] = torch.from_numpy(f[np.ix_(o, o)].T.reshape(-1).copy()).to(dev)
    out["FVO"] = torch.from_numpy(f[np.ix_(vv, o)].T.reshape(-1).copy()).to(dev)
    out["FVV"] = torch.from_numpy(f[np.ix_(vv, vv)].T.reshape(-1).copy()).to(dev)

    def t(x):
        return torch.from_numpy(np.ascontiguousarray(x)).to(dev)

    ip, jp = pair_index(no)            # I > J
    ap, bp = pair_index(nv)            # A > B
    ak_a = np.tile(np.arange(nv), no) + no      # VO pair (A,K): A fastest
    ak_k = np.repeat(np.arange(no), nv)
    kb = kblock or max(1, min(4096, (1 << 27) // max(1, rank * max(nv, no) * 4)))
    spec = {
        "OOOO": (ip, jp, ip, jp),
        "VOOO": (ak_a, ak_k, ip, jp),
        "VVOO": (ap + no, bp + no, ip, jp),
        "VOVO": (ak_a, ak_k, ak_a, ak_k),
        "VOVV": (ak_a, ak_k, ap + no, bp + no),
        "VVVV": (ap + no, bp + no, ap + no, bp + no),
    }
    for name in classes:
        pr, qr, rc, sc = spec[name]
        if len(pr) == 0 or len(rc) == 0:
            out[name] = torch.zeros(len(pr) * len(rc), dtype=torch.complex128, device=dev)
            continue
        out[name] = _build_class(B, t(pr), t(qr), t(rc), t(sc), True, dev, kb)
    return out


def hermitian_perturbation(n: int, seed: int, sigma: float = 0.01, irreps=None) -> np.ndarray:
    """Seeded Hermitian one-electron perturbation V (n x n, spinors occupied
    then virtual): V_pq (p < q) with Re, Im ~ N(0, sigma^2), V_qp its
    conjugate, V_pp ~ N(0, sigma^2).  irreps (the irrep of each spinor):
    V_pq = 0 between spinors of different irreps (a totally symmetric
    operator)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    re = rng.normal(0.0, sigma, (n, n))
    im = rng.normal(0.0, sigma, (n, n))
    up = np.triu(re + 1j * im, 1)
    v = up + up.conj().T + np.diag(rng.normal(0.0, sigma, n))
    if irreps is not None:
        g = np.asarray(irreps)
        v[g[:, None] != g[None, :]] = 0.0
    return v


def _pair_c1(x, y, n):
    return y * (2 * n - y - 1) // 2 + (x - y - 1)


def block_maps(sym):
    """SYMTAB's symmetry-blocked orderings (ccdriv.F:1050-1187) for the
    Z2^k double group of make_system(sym=...): for each kind ('oo', 'vv',
    'vo', 'occ', 'vir'), one array per block of C1 indices (packed pair,
    VO pair K*nv+A, or spinor).  Pair sub-blocks (i >= j, pair irrep c) are
    ordered by j, J outer and I inner; VO sub-blocks by the occupied irrep,
    K outer and A inner."""
    no_l, nv_l = (list(map(int, x)) for x in sym)
    n = len(no_l)
    io = np.concatenate([[0], np.cumsum(no_l)])
    iv = np.concatenate([[0], np.cumsum(nv_l)])
    NO, NV = int(io[-1]), int(iv[-1])
    g = np.arange(n)
    ptab = np.bitwise_xor.outer(g, g)
    partner = np.zeros((n, n), dtype=np.int64)      # partner[j, c]: i with ptab[i, j] = c
    for i in range(n):
        for j in range(n):
            partner[j, ptab[i, j]] = i
    maps = {"oo": [], "vv": [], "vo": [], "occ": [], "vir": []}
    for c in range(n):
        for key, cnt, off, tot in (("oo", no_l, io, NO), ("vv", nv_l, iv, NV)):
            out = []
            for jr in range(n):
                ir = partner[jr, c]
                if ir < jr:
                    continue
                for j in range(cnt[jr]):
                    lo = j + 1 if ir == jr else 0
                    i = np.arange(lo, cnt[ir])
                    out.append(_pair_c1(off[ir] + i, off[jr] + j, tot))
            maps[key].append(np.concatenate(out).astype(np.int64) if out else
                             np.zeros(0, np.int64))
        vo = []
        for kr in range(n):
            ar = partner[kr, c]
            k = np.arange(no_l[kr])[:, None]
            a = np.arange(nv_l[ar])[None, :]
            vo.append(((io[kr] + k) * NV + iv[ar] + a).reshape(-1))
        maps["vo"].append(np.concatenate(vo).astype(np.int64))
        maps["occ"].append(np.arange(io[c], io[c + 1]))
        maps["vir"].append(np.arange(iv[c], iv[c + 1]))
    return maps


_BLOCK_KINDS = {"FOO": ("occ", "occ"), "FVO": ("vir", "occ"), "T1": ("vir", "occ"),
                "FVV": ("vir", "vir"), "OOOO": ("oo", "oo"), "VOOO": ("vo", "oo"),
                "VVOO": ("vv", "oo"), "T2": ("vv", "oo"), "VOVO": ("vo", "vo"),
                "VOVV": ("vo", "vv"), "VVVV": ("vv", "vv")}


def to_blocks(name, x, sym, maps=None):
    """C1 array of class / block kind `name` (flat, column-major, numpy or
    torch) -> relccsd's symmetry-blocked layout: the blocks of every compound
    irrep (or irrep), each column-major, concatenated."""
    maps = maps or block_maps(sym)
    rk, ck = _BLOCK_KINDS[name]
    NO, NV = sum(map(int, sym[0])), sum(map(int, sym[1]))
    nrows = {"oo": NO * (NO - 1) // 2, "vv": NV * (NV - 1) // 2, "vo": NO * NV,
             "occ": NO, "vir": NV}[rk]
    if isinstance(x, torch.Tensor):
        X = x.reshape(-1, nrows) if nrows else x.reshape(0, 0)
        parts = []
        for r, c in zip(maps[rk], maps[ck]):
            if len(r) and len(c):
                rt = torch.as_tensor(r, device=x.device)
                ct = torch.as_tensor(c, device=x.device)
                parts.append(X.index_select(0, ct).index_select(1, rt).reshape(-1))
        return torch.cat(parts) if parts else x.new_zeros(0)
    X = np.asarray(x).reshape(-1, nrows) if nrows else np.asarray(x).reshape(0, 0)
    parts = [X[np.ix_(c, r)].reshape(-1) for r, c in zip(maps[rk], maps[ck]) if len(r) and len(c)]
    return np.concatenate(parts) if parts else np.zeros(0, np.asarray(x).dtype)


def batch_sizes(nsys: int, nrep: int, seed: int) -> list:
    """Sizes of the batched many-systems workload: per system, occupied ~
    U{4..16} and virtual ~ U{40..160}, both split over nrep irreps with one
    Dirichlet(2) draw of proportions; returns (o, v, o per irrep, v per
    irrep, generator seed) tuples."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(nsys):
        o, v = int(rng.integers(4, 17)), int(rng.integers(40, 161))
        p = rng.dirichlet([2.0] * nrep)
        out.append((o, v, dirichlet_split(o, p), dirichlet_split(v, p), 1000 + k))
    return out


def dirichlet_split(total: int, props) -> list:
    """Integer counts summing to `total` in the given proportions (largest
    remainders)."""
    raw = np.asarray(props, dtype=np.float64) * total
    out = np.floor(raw).astype(int)
    for i in np.argsort(-(raw - out))[: total - int(out.sum())]:
        out[i] += 1
    return out.tolist()
