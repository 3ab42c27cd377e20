"""MOLFDIR integral files: writers for MRCONEE and MDCINT and a seeded
synthetic Kramers-reduced integral list.

Record layouts (Fortran unformatted sequential, 4-byte record markers) as
relccsd reads them:

MRCONEE (RDSYMI ccdriv.F:2978-2988, GETFM ccints.F:1091-1096)
  1  NSP, BREIT (logical*4), ECORE (real*8)
  2  NSYMRP, names (character*14 each), the occupied count per irrep (not read)
  3  NREP, names of the 2 NREP abelian irreps (character*4 each)
  4  MULTB(1:2NREP, 1:2NREP, 1), column-major
  5  (higher-group irrep, abelian irrep, orbital energy) per spinor
  6  the one-electron (Fock) matrix, (Re, Im) per element, NSP x NSP column-major

MDPROP (RDPROP ccints.F:569-764): per operator a label record (character*32:
'********', date (8), storage info (8), name (8)) followed by the operator
matrix, (Re, Im) per element, NSP x NSP column-major.

MDCINT (RDINTS ccints.F:117-118, GET2INTS ccints.F:819-835)
  header  DATEX (character*10), TIMEX (character*8), NKR, (KR(i), KR(-i), i = 1..NKR)
  records IKR, JKR, NZ, (INDK, INDL) x NZ, (Re, Im) x NZ  -- the integrals
          (IKR JKR | INDK INDL) over Kramers pairs (negative: the barred
          partner), Mulliken order
  end     IKR = JKR = 0
  continuation (GET2INTS ccints.F:832-838): IKR = 0, JKR != 0, then a record
          naming the next file (CHARACTER*8), which holds integral records
          only
"""
from __future__ import annotations

import struct

import numpy as np


def _record(f, payload: bytes):
    n = len(payload)
    f.write(struct.pack("<i", n))
    f.write(payload)
    f.write(struct.pack("<i", n))


def write_mrconee(path, eps, irreps, nelec, h1, multb=None, ecore=0.0, breit=False):
    """irreps: 1-based fermion irrep of each spinor; nelec: occupied count per
    irrep; h1: NSP x NSP complex one-electron matrix; multb: 2NREP x 2NREP
    table (default the C1 one)."""
    eps = np.asarray(eps, dtype=np.float64)
    nsp = eps.size
    irreps = np.asarray(irreps, dtype=np.int32)
    nelec = np.asarray(nelec, dtype=np.int32)
    nrep = nelec.size
    if multb is None:
        multb = np.array([[2, 1], [1, 2]], dtype=np.int32)
    multb = np.asarray(multb, dtype=np.int32)
    with open(path, "wb") as f:
        _record(f, struct.pack("<iid", nsp, 1 if breit else 0, ecore))
        names = b"".join(f"A{i + 1:02d}".ljust(14).encode() for i in range(nrep))
        _record(f, struct.pack("<i", nrep) + names + nelec.tobytes())
        rn = b"".join(f" A{i + 1:02d}".encode() for i in range(nrep)) + \
            b"".join(f" a{i + 1:02d}".encode() for i in range(nrep))
        _record(f, struct.pack("<i", nrep) + rn)
        _record(f, multb.T.astype(np.int32).tobytes())            # column-major
        rec5 = np.zeros(nsp, dtype=[("g", "<i4"), ("a", "<i4"), ("e", "<f8")])
        rec5["g"], rec5["a"], rec5["e"] = irreps, irreps, eps
        _record(f, rec5.tobytes())
        h = np.asarray(h1, dtype=np.complex128).reshape(nsp, nsp)
        _record(f, np.ascontiguousarray(h.T).reshape(-1).view(np.float64).tobytes())


def write_mdprop(path, operators):
    """operators: {name (at most 8 characters): NSP x NSP complex matrix}."""
    with open(path, "wb") as f:
        for name, mat in operators.items():
            label = ("********" + "01Jan26 " + "SQUARE  " + name.ljust(8)[:8]).encode()
            _record(f, label)
            m = np.asarray(mat, dtype=np.complex128)
            _record(f, np.ascontiguousarray(m.T).reshape(-1).view(np.float64).tobytes())


def write_mdcint(path, kr, ikr, jkr, kkr, lkr, val, parts: int = 1):
    """kr: (NKR, 2) spinor indices (1-based) of the unbarred / barred member
    of each Kramers pair.  The canonical integrals (ikr, jkr | kkr, lkr) are
    written one record per run of equal (ikr, jkr), in the given order.
    parts > 1 spreads the records over that many files, as GET2INTS reads
    them (ccints.F:832-838): each file but the last ends with the
    continuation record (IKR = 0, JKR = 1, NZ = 0) and a record naming the
    next file (CHARACTER*8: MDCINT1, MDCINT2, .. beside path), which holds
    integral records only."""
    import os
    kr = np.asarray(kr, dtype=np.int32)
    ikr, jkr = np.asarray(ikr, np.int32), np.asarray(jkr, np.int32)
    kkr, lkr = np.asarray(kkr, np.int32), np.asarray(lkr, np.int32)
    val = np.asarray(val, dtype=np.complex128)
    recs = []
    n = ikr.size
    s = 0
    while s < n:
        e = s + 1
        while e < n and ikr[e] == ikr[s] and jkr[e] == jkr[s]:
            e += 1
        kl = np.stack([kkr[s:e], lkr[s:e]], axis=1).astype("<i4")
        recs.append(struct.pack("<iii", int(ikr[s]), int(jkr[s]), e - s) + kl.tobytes() +
                    val[s:e].view(np.float64).astype("<f8").tobytes())
        s = e
    parts = max(1, min(int(parts), max(len(recs), 1)))
    cuts = [round(i * len(recs) / parts) for i in range(parts + 1)]
    folder = os.path.dirname(os.path.abspath(path))
    for part in range(parts):
        name = path if part == 0 else os.path.join(folder, f"MDCINT{part}")
        with open(name, "wb") as f:
            if part == 0:
                _record(f, b"01Jan2026 " + b"00:00:00" + struct.pack("<i", kr.shape[0]) +
                        kr.astype("<i4").tobytes())
            for rec in recs[cuts[part]:cuts[part + 1]]:
                _record(f, rec)
            if part + 1 < parts:
                _record(f, struct.pack("<iii", 0, 1, 0))
                _record(f, f"MDCINT{part + 1}".ljust(8).encode())
            else:
                _record(f, struct.pack("<iii", 0, 0, 0))


def synthetic_list(nkr: int, seed: int, keep: float = 0.7, scale: float = 0.05):
    """A seeded Kramers-reduced list: records (IKR, JKR) for IKR = 1..NKR and
    0 < |JKR| <= IKR; in each, (KKR, LKR) with 0 < KKR <= IKR, 0 < |LKR| <= KKR,
    each kept with probability `keep`; complex values N(0, scale^2).  The
    list is not restricted to one representative per symmetry orbit: the
    expansion then writes some sorted elements more than once, and the
    last write (in file order) is the one relccsd keeps."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ik, jk, kk, lk = [], [], [], []
    for i in range(1, nkr + 1):
        for j in list(range(-i, 0)) + list(range(1, i + 1)):
            for k in range(1, i + 1):
                for l in list(range(-k, 0)) + list(range(1, k + 1)):
                    ik.append(i)
                    jk.append(j)
                    kk.append(k)
                    lk.append(l)
    ik, jk, kk, lk = (np.asarray(x, np.int32) for x in (ik, jk, kk, lk))
    m = rng.random(ik.size) < keep
    ik, jk, kk, lk = ik[m], jk[m], kk[m], lk[m]
    val = scale * (rng.normal(size=ik.size) + 1j * rng.normal(size=ik.size))
    return ik, jk, kk, lk, val


def synthetic_files(dirpath, nkr: int, nocc: int, seed: int, keep: float = 0.7, parts: int = 1,
                    breit: bool = False, scale: float = 0.05):
    """C1 MRCONEE + MDCINT in dirpath: NSP = 2 NKR spinors with shuffled
    Kramers pairs, the first nocc spinors (input order) occupied, orbital
    energies ascending, a Hermitian one-electron matrix (diagonal eps,
    off-diagonal N(0, 0.02^2)), integral values N(0, scale^2).  breit:
    MRCONEE's BREIT flag set and a second seeded list (values N(0, 0.01^2))
    written as MDBINT.  Returns the generated arrays."""
    import os
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    nsp = 2 * nkr
    perm = rng.permutation(nsp) + 1
    kr = perm.reshape(nkr, 2)
    eps = np.sort(np.concatenate([rng.uniform(-2.0, -0.3, nocc), rng.uniform(0.2, 5.0, nsp - nocc)]))
    h1 = 0.02 * (rng.normal(size=(nsp, nsp)) + 1j * rng.normal(size=(nsp, nsp)))
    h1 = 0.5 * (h1 + h1.conj().T)
    h1[np.arange(nsp), np.arange(nsp)] = eps
    ik, jk, kk, lk, val = synthetic_list(nkr, seed, keep, scale)
    write_mrconee(os.path.join(dirpath, "MRCONEE"), eps, np.ones(nsp, np.int32), [nocc], h1,
                  breit=breit)
    write_mdcint(os.path.join(dirpath, "MDCINT"), kr, ik, jk, kk, lk, val, parts=parts)
    out = {"nsp": nsp, "nocc": nocc, "kr": kr, "eps": eps, "h1": h1, "ikr": ik, "jkr": jk,
           "kkr": kk, "lkr": lk, "val": val}
    if breit:
        bl = synthetic_list(nkr, seed + 1000, keep, scale=0.01)
        write_mdcint(os.path.join(dirpath, "MDBINT"), kr, *bl)
        out["breit"] = dict(zip(("ikr", "jkr", "kkr", "lkr", "val"), bl))
    return out


def abelian_multb(nrep: int) -> np.ndarray:
    """MULTB(1:2NREP, 1:2NREP, 1) of the abelian double group Z2^k x
    {fermion, boson} (fermion irreps 1..NREP, boson irreps NREP+1..2NREP):
    irreps multiply by xor of their 0-based index within each half, and two
    fermion (or two boson) irreps give a boson irrep."""
    out = np.zeros((2 * nrep, 2 * nrep), dtype=np.int32)
    for i in range(2 * nrep):
        for j in range(2 * nrep):
            g = (i % nrep) ^ (j % nrep)
            fermion = (i < nrep) != (j < nrep)
            out[i, j] = (1 + g) if fermion else (nrep + 1 + g)
    return out


def synthetic_files_sym(dirpath, nkr_irrep, nocc_irrep, seed: int, keep: float = 0.7,
                        breit: bool = False, scale: float = 0.05):
    """MRCONEE + MDCINT with the symmetry of Z2^k x {fermion, boson}
    (len(nkr_irrep) = 2, 4 or 8 fermion irreps).  Both members of a Kramers
    pair lie in the pair's irrep; spinors are written in a shuffled order, and
    within each irrep the first nocc_irrep[r] of them (file order) are
    occupied; the one-electron matrix is irrep-diagonal; the canonical list
    keeps only integrals whose four Kramers pairs have irreps that multiply
    to the totally symmetric one (so every image is allowed).  breit:
    MRCONEE's BREIT flag set and a second list, filtered the same way, as
    MDBINT.  Integral values N(0, scale^2)."""
    import os
    nrep = len(nkr_irrep)
    rng = np.random.Generator(np.random.PCG64(seed + 104729))
    pair_irrep = np.concatenate([np.full(n, r, dtype=np.int64) for r, n in enumerate(nkr_irrep)])
    nkr = pair_irrep.size
    nsp = 2 * nkr
    perm = rng.permutation(nsp) + 1
    kr = perm.reshape(nkr, 2)
    irreps = np.zeros(nsp, dtype=np.int32)
    irreps[kr[:, 0] - 1] = pair_irrep + 1
    irreps[kr[:, 1] - 1] = pair_irrep + 1
    occ = np.zeros(nsp, dtype=bool)
    for r in range(nrep):
        members = np.nonzero(irreps == r + 1)[0]
        occ[members[:nocc_irrep[r]]] = True
    eps = np.where(occ, rng.uniform(-2.0, -0.3, nsp), rng.uniform(0.2, 5.0, nsp))
    h1 = 0.02 * (rng.normal(size=(nsp, nsp)) + 1j * rng.normal(size=(nsp, nsp)))
    h1 = 0.5 * (h1 + h1.conj().T)
    h1[irreps[:, None] != irreps[None, :]] = 0.0
    h1[np.arange(nsp), np.arange(nsp)] = eps
    ik, jk, kk, lk, val = synthetic_list(nkr, seed, keep, scale)
    g = pair_irrep
    ok = (g[np.abs(ik) - 1] ^ g[np.abs(jk) - 1] ^ g[np.abs(kk) - 1] ^ g[np.abs(lk) - 1]) == 0
    ik, jk, kk, lk, val = ik[ok], jk[ok], kk[ok], lk[ok], val[ok]
    nelec = np.array(nocc_irrep, dtype=np.int32)
    write_mrconee(os.path.join(dirpath, "MRCONEE"), eps, irreps, nelec, h1,
                  multb=abelian_multb(nrep), breit=breit)
    write_mdcint(os.path.join(dirpath, "MDCINT"), kr, ik, jk, kk, lk, val)
    if breit:
        # a second list with the same symmetry filter, as MDBINT
        bl = synthetic_list(nkr, seed + 1000, keep, scale=0.01)
        okb = (g[np.abs(bl[0]) - 1] ^ g[np.abs(bl[1]) - 1] ^ g[np.abs(bl[2]) - 1] ^
               g[np.abs(bl[3]) - 1]) == 0
        write_mdcint(os.path.join(dirpath, "MDBINT"), kr, *(x[okb] for x in bl))
    no_i = [int(x) for x in nelec]
    nv_i = [int(np.sum(irreps == r + 1)) - no_i[r] for r in range(nrep)]
    return {"nsp": nsp, "irreps": irreps, "nelec": nelec, "no": no_i, "nv": nv_i, "kr": kr,
            "eps": eps, "h1": h1, "ikr": ik, "jkr": jk, "kkr": kk, "lkr": lk, "val": val}
