"""CPU restatement of relccsd's integral front end, single reference, C1 and
abelian double groups (test infrastructure only): MKINDEX (ccints.F:1421-1580), GET2INTS
(ccints.F:765-1029, the non-Fock-space branch), the class sort of RDINTS
(ccints.F:33-569) and the Fock build of GETFM / MAKEFM
(ccints.F:1031-1419).

* Spinors: within each irrep, in input (MRCONEE) order, the first NELEC are
  occupied, the rest virtual (SYMTAB ccdriv.F:832-858).  Positions are those
  of the C1 working order: occupied irrep-major (file order inside an irrep),
  then virtual irrep-major -- for one irrep, the order among the occupied /
  virtual spinors.  With symmetry the classes come out in that C1 order and
  SymLayout.to_sym gives relccsd's symmetry-blocked layout.
* MKINDEX: type of a spinor pair (I, J) = 1 oo, 2 vo, 3 ov, 4 vv (I, J
  occupied or virtual); packed pair index of oo / vv pairs for I > J in
  position order, of vo pairs (A fastest) for all.
* GET2INTS expands each canonical (IKR JKR | KKR LKR) into 16 index
  quadruples: time reversal of the (IJ) pair (KR(-JKR), KR(-IKR), sign
  sign(IKR JKR)), of the (KL) pair, particle interchange (IJ) <-> (KL), and
  complex conjugation (I <-> J, K <-> L, imaginary part negated).  Each
  (ij|kl) = <ik|jl> goes to the class of (ik; jl) -- the direct part when
  (i, k) is a canonical pair, the exchange part at (k, i) otherwise.
* RDINTS stores every expanded value by assignment, in file order, so the
  last write of an element wins; class = direct - exchange.
* With BREIT, the Breit list (MDBINT) is expanded with ITRSGN = -1 (the
  signs of the time-reversed pair images flipped), sorted the same way, and
  the Coulomb classes are added to it.
* GETFM: F = conj(h) of MRCONEE record 6 on the oo / vo / vv blocks, then
  MAKEFM adds the occupied two-electron sums, E0 = sum Re F_ii,
  E1 = -1/2 sum_ik Re <ik||ik>, and the orbital energies become Re diag F.

Pinned by tests/test_oracle_rdints.py against the reference's own sort of
seeded MRCONEE / MDCINT files (tests/golden/sort_*.npz and symsort_*.npz,
the latter with 2, 4 and 8 irreps).
"""
from __future__ import annotations

import numpy as np

from molfdir_amd.synth import pair_index


def _pidx(p, q, n):
    return q * n - q * (q + 1) // 2 + (p - q - 1)


class Spinors:
    """Occupation, C1 position and MKINDEX tables of an NSP-spinor set."""

    def __init__(self, irreps, nelec):
        irreps = np.asarray(irreps)
        self.nsp = irreps.size
        self.occ = np.zeros(self.nsp, dtype=bool)
        seen = {}
        for s in range(self.nsp):
            r = int(irreps[s])
            seen[r] = seen.get(r, 0) + 1
            self.occ[s] = seen[r] <= int(nelec[r - 1])
        self.pos = np.zeros(self.nsp, dtype=np.int64)
        for flag in (True, False):
            sel = np.nonzero(self.occ == flag)[0]
            order = sel[np.argsort(irreps[sel], kind="stable")]   # irrep-major, file order
            self.pos[order] = np.arange(order.size)
        self.no, self.nv = int(self.occ.sum()), int((~self.occ).sum())

    def ptype(self, i, j):
        """MKINDEX INDEX(I,J,1) for 0-based spinors."""
        return (1 if self.occ[i] else 2) + (0 if self.occ[j] else 2)

    def pindex(self, i, j):
        """MKINDEX INDEX(I,J,4) (1-based, 0 = not a stored pair)."""
        t = self.ptype(i, j)
        pi, pj = int(self.pos[i]), int(self.pos[j])
        if t == 1:
            return _pidx(pi, pj, self.no) + 1 if pi > pj else 0
        if t == 4:
            return _pidx(pi, pj, self.nv) + 1 if pi > pj else 0
        if t == 2:
            return pj * self.nv + pi + 1
        return 0


CLASS_OF_TYPE = {1: "OOOO", 2: "VOOO", 4: "VVOO", 6: "VOVO", 14: "VOVV", 16: "VVVV"}


def expand(sp: Spinors, kr, ikr, jkr, kkr, lkr, itrsgn=1):
    """GET2INTS: for each canonical integral, the list of (class type, row,
    col, re sign, im sign) of its stored expansions, in the order RDINTS
    assigns them.  kr[(i)] / kr[(-i)]: 0-based spinor of Kramers pair i.
    itrsgn: ITRSGN, multiplying the signs of the time-reversed pair images
    (ccints.F:855, 939): 1 for the Coulomb list, -1 for the Breit list."""
    out = []
    for n in range(len(ikr)):
        IK, JK, KK, LK = int(ikr[n]), int(jkr[n]), int(kkr[n]), int(lkr[n])
        ijsign = int(np.sign(IK * JK)) * itrsgn
        i0, j0 = kr[IK], kr[JK]
        ijsg = 1
        for _ in range(2):
            k0, l0 = kr[KK], kr[LK]
            klsg = 1
            for _ in range(2):
                i1, j1, k1, l1 = i0, j0, k0, l0
                for _ in range(2):
                    i, j, k, l = i1, j1, k1, l1
                    for icc in (1, -1):
                        jl = sp.pindex(j, l)
                        if jl != 0 and i != k:
                            ik = sp.pindex(i, k)
                            if ik != 0:
                                t = (sp.ptype(j, l) - 1) * 4 + sp.ptype(i, k)
                                out.append((n, t, ik - 1, jl - 1, ijsg * klsg, icc * ijsg * klsg))
                            else:
                                t = (sp.ptype(j, l) - 1) * 4 + sp.ptype(k, i)
                                out.append((n, -t, sp.pindex(k, i) - 1, jl - 1, ijsg * klsg,
                                            icc * ijsg * klsg))
                        i, j, k, l = j1, i1, l1, k1
                    i1, j1, k1, l1 = k0, l0, i0, j0
                k0, l0 = kr[-LK], kr[-KK]
                klsg = int(np.sign(KK * LK)) * itrsgn
            i0, j0 = kr[-JK], kr[-IK]
            ijsg = ijsign
    return out


def class_shapes(no, nv):
    noot, nvvt, nvo = no * (no - 1) // 2, nv * (nv - 1) // 2, no * nv
    return {"OOOO": (noot, noot), "VOOO": (nvo, noot), "VVOO": (nvvt, noot),
            "VOVO": (nvo, nvo), "VOVV": (nvo, nvvt), "VVVV": (nvvt, nvvt)}


def sort_classes(sp: Spinors, kr, ikr, jkr, kkr, lkr, val, itrsgn=1):
    """RDINTS's six classes (packed C1 layout, rows fastest)."""
    shapes = class_shapes(sp.no, sp.nv)
    direct = {c: np.zeros(r * q, np.complex128) for c, (r, q) in shapes.items()}
    exch = {c: np.zeros(r * q, np.complex128) for c, (r, q) in shapes.items()}
    for n, t, row, col, s_re, s_im in expand(sp, kr, ikr, jkr, kkr, lkr, itrsgn):
        c = CLASS_OF_TYPE.get(abs(t))
        if c is None:
            continue
        rows = shapes[c][0]
        v = val[n]
        (direct if t > 0 else exch)[c][row + col * rows] = complex(v.real * s_re, v.imag * s_im)
    return {c: direct[c] - exch[c] for c in shapes}


def sort_classes_breit(sp: Spinors, kr, coulomb, breit):
    """RDINTS with BREIT (ccints.F:230-238, 354-358, 457-461, 496-505): the
    Breit list sorted like the Coulomb one with ITRSGN = -1, then the Coulomb
    classes added to it.  coulomb, breit: (ikr, jkr, kkr, lkr, val)."""
    c = sort_classes(sp, kr, *coulomb)
    b = sort_classes(sp, kr, *breit, itrsgn=-1)
    return {k: b[k] + c[k] for k in c}


def kr_table(kr_pairs):
    """{i: spinor of KR(i), -i: spinor of KR(-i)} (0-based spinors) from the
    (NKR, 2) table of 1-based indices written in the MDCINT header."""
    kr = {}
    for i, (a, b) in enumerate(np.asarray(kr_pairs), start=1):
        kr[i] = int(a) - 1
        kr[-i] = int(b) - 1
    return kr


def fock(sp: Spinors, h1, classes, eps_file=None, useoe=False, norecmp=False):
    """GETFM + MAKEFM: FOO, FVO, FVV (packed column-major), E0, E1, eps.
    eps_file: MRCONEE record 5's orbital energy of every spinor (file order),
    used with namelist CCSORT's USEOE (eps kept, ccints.F:1260, 1287) and
    NORECMP (F = diag(eps_file), FVO = 0, E0 = sum of the occupied ones, E1
    from the integrals as always, ccints.F:1161-1198)."""
    from .ccsd_oracle import Layout
    no, nv = sp.no, sp.nv
    occ_s = np.nonzero(sp.occ)[0][np.argsort(sp.pos[sp.occ])]
    vir_s = np.nonzero(~sp.occ)[0][np.argsort(sp.pos[~sp.occ])]
    if useoe or norecmp:
        ef = np.asarray(eps_file, dtype=np.float64)
        eps_w = np.concatenate([ef[occ_s], ef[vir_s]])
    if norecmp:
        L = Layout(no, nv)
        oooo = L.unpack_pp(classes["OOOO"], no, no)
        foo = np.diag(eps_w[:no]).astype(np.complex128)
        fvv = np.diag(eps_w[no:]).astype(np.complex128)
        fvo = np.zeros((nv, no), np.complex128)
        return {"FOO": foo.T.reshape(-1), "FVO": fvo.T.reshape(-1), "FVV": fvv.T.reshape(-1),
                "E0": float(np.sum(eps_w[:no])),
                "E1": -0.5 * float(np.real(np.einsum("ikik->", oooo))), "eps": eps_w}
    h = np.asarray(h1)
    foo = np.conj(h[np.ix_(occ_s, occ_s)])
    fvo = np.conj(h[np.ix_(vir_s, occ_s)])
    fvv = np.conj(h[np.ix_(vir_s, vir_s)])
    L = Layout(no, nv)
    oooo = L.unpack_pp(classes["OOOO"], no, no)        # <kl||ij>
    vooo = L.unpack_vo_pp(classes["VOOO"], no)          # <ak||ij>
    vovo = L.unpack_vovo(classes["VOVO"])               # <ak||ci>
    foo = foo + np.einsum("ikjk->ij", oooo)
    fvo = fvo + np.einsum("akik->ai", vooo)
    fvv = fvv + np.einsum("akck->ac", vovo)
    e1 = -0.5 * float(np.real(np.einsum("ikik->", oooo)))
    e0 = float(np.real(np.trace(foo)))
    eps = np.concatenate([np.real(np.diag(foo)), np.real(np.diag(fvv))])
    if useoe:
        eps = eps_w
    return {"FOO": foo.T.reshape(-1), "FVO": fvo.T.reshape(-1), "FVV": fvv.T.reshape(-1),
            "E0": e0, "E1": e1, "eps": eps}
