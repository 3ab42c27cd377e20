"""Restatement of relccsd's symmetry-blocked layouts (test infrastructure).

SYMTAB (ccdriv.F:746-1425) orders every packed quantity by compound irrep of
the abelian double group.  This module reproduces those orderings as index
maps into the C1 layouts used by molfdir_amd.synth / oracle.ccsd_oracle, so a
system drawn in C1 form with a symmetry pattern can be written in relccsd's
own symmetry layout (for the reference harness) and read back.

Conventions (all 0-based here):
* fermion irreps 0..nrep-1; pair irreps (bosons) 0..nrep-1, 0 the totally
  symmetric one.  ``ptab[i, j]`` is the pair irrep of fermion irreps i, j:
  MULTB(I,J,1) after RDSYMI drops the boson offset (ccdriv.F:2996-3020).
* global spinor order: occupied irrep-major, then virtual irrep-major -- the
  order SYMTAB puts the orbital energies in (ccdriv.F:1392-1402) and the C1
  index of every spinor here.
* packed pairs of compound irrep c: sub-blocks (i, j) with i >= j and
  ptab[i, j] = c, ordered by j ascending (IIOOT/IIVVT, ccdriv.F:1071-1082);
  inside a diagonal sub-block J outer, I > J inner; off-diagonal J outer,
  I inner (DENOM cceqns.F:2408-2448).
* VO pairs (A, K) of compound c: sub-blocks (a_rep, k_rep) with
  ptab[a_rep, k_rep] = c ordered by k_rep ascending (IIVO, ccdriv.F:1050-
  1057), inside K outer, A inner.
* class blocks one per compound irrep, consecutive (IOOOOTT .. IVVVVTT,
  ccdriv.F:1163-1187); Fock and T1 blocks one per irrep (IOO/IVO/IVV).
"""
from __future__ import annotations

import numpy as np


def xor_table(nrep: int) -> np.ndarray:
    """Pair-irrep table of the abelian double group Z2^k x {fermion, boson}
    (nrep = 2^k fermion irreps): irrep products add their Z2^k labels."""
    assert nrep >= 1 and nrep & (nrep - 1) == 0, "nrep must be a power of two"
    g = np.arange(nrep)
    return np.bitwise_xor.outer(g, g)


def multb_full(nrep: int) -> np.ndarray:
    """The 2 nrep x 2 nrep MULTB(.,.,1) that MRCONEE holds for xor_table
    (1-based: fermions 1..nrep, bosons nrep+1..2 nrep), as RDSYMI reads it
    (ccdriv.F:2985)."""
    out = np.zeros((2 * nrep, 2 * nrep), dtype=np.int32)
    for i in range(2 * nrep):
        for j in range(2 * nrep):
            gi, fi = i % nrep, int(i < nrep)
            gj, fj = j % nrep, int(j < nrep)
            g, f = gi ^ gj, fi ^ fj
            out[i, j] = (1 + g) if f else (nrep + 1 + g)
    return out


class SymLayout:
    def __init__(self, no, nv, ptab=None):
        self.no = np.asarray(no, dtype=np.int64)
        self.nv = np.asarray(nv, dtype=np.int64)
        self.nrep = len(self.no)
        self.ptab = xor_table(self.nrep) if ptab is None else np.asarray(ptab)
        n = self.nrep
        self.NO, self.NV = int(self.no.sum()), int(self.nv.sum())
        self.io = np.concatenate([[0], np.cumsum(self.no)])[:-1]
        self.iv = np.concatenate([[0], np.cumsum(self.nv)])[:-1]
        # partner: the i with ptab[i, j] = c
        self.partner = np.full((n, n), -1, dtype=np.int64)
        for i in range(n):
            for j in range(n):
                self.partner[j, self.ptab[i, j]] = i
        self.oo = [self._pairs(self.no, self.io, self.NO, c) for c in range(n)]
        self.vv = [self._pairs(self.nv, self.iv, self.NV, c) for c in range(n)]
        self.vo = [self._vo(c) for c in range(n)]

    @staticmethod
    def _pair_c1(x, y, n):
        """C1 packed index of (x > y): y outer, x inner."""
        return y * (2 * n - y - 1) // 2 + (x - y - 1)

    def _pairs(self, cnt, off, ntot, c):
        out = []
        for jr in range(self.nrep):
            ir = self.partner[jr, c]
            if ir < jr:
                continue
            for j in range(cnt[jr]):
                i0 = j + 1 if ir == jr else 0
                for i in range(i0, cnt[ir]):
                    out.append(self._pair_c1(off[ir] + i, off[jr] + j, ntot))
        return np.asarray(out, dtype=np.int64)

    def _vo(self, c):
        out = []
        for kr in range(self.nrep):
            ar = self.partner[kr, c]
            for k in range(self.no[kr]):
                for a in range(self.nv[ar]):
                    out.append((self.io[kr] + k) * self.NV + self.iv[ar] + a)
        return np.asarray(out, dtype=np.int64)

    # ---- class descriptions: (row maps, col maps, C1 leading dimension)
    def blocks(self, name):
        """Per compound irrep: (rowmap, colmap) into the C1 array of `name`."""
        o = [np.arange(self.no[i]) + self.io[i] for i in range(self.nrep)]
        v = [np.arange(self.nv[i]) + self.iv[i] for i in range(self.nrep)]
        if name == "FOO":
            return [(o[i], o[i]) for i in range(self.nrep)]
        if name in ("FVO", "T1"):
            return [(v[i], o[i]) for i in range(self.nrep)]
        if name == "FVV":
            return [(v[i], v[i]) for i in range(self.nrep)]
        rows, cols = {
            "OOOO": (self.oo, self.oo), "VOOO": (self.vo, self.oo),
            "VVOO": (self.vv, self.oo), "T2": (self.vv, self.oo),
            "VOVO": (self.vo, self.vo), "VOVV": (self.vo, self.vv),
            "VVVV": (self.vv, self.vv),
        }[name]
        return list(zip(rows, cols))

    def c1_shape(self, name):
        no, nv = self.NO, self.NV
        noot, nvvt, nvo = no * (no - 1) // 2, nv * (nv - 1) // 2, no * nv
        return {"FOO": (no, no), "FVO": (nv, no), "T1": (nv, no), "FVV": (nv, nv),
                "OOOO": (noot, noot), "VOOO": (nvo, noot), "VVOO": (nvvt, noot),
                "T2": (nvvt, noot), "VOVO": (nvo, nvo), "VOVV": (nvo, nvvt),
                "VVVV": (nvvt, nvvt)}[name]

    def size(self, name):
        return int(sum(len(r) * len(c) for r, c in self.blocks(name)))

    def to_sym(self, name, x):
        """C1 flat column-major array -> concatenated symmetry blocks."""
        x = np.asarray(x).reshape(self.c1_shape(name)[::-1]).T
        parts = [x[np.ix_(r, c)].T.reshape(-1) for r, c in self.blocks(name)]
        return np.concatenate(parts) if parts else np.zeros(0, x.dtype)

    def to_c1(self, name, s):
        """Symmetry blocks -> C1 flat array (zero off the symmetry blocks)."""
        shape = self.c1_shape(name)
        out = np.zeros(shape, dtype=np.asarray(s).dtype)
        pos = 0
        for r, c in self.blocks(name):
            n = len(r) * len(c)
            out[np.ix_(r, c)] = np.asarray(s[pos:pos + n]).reshape(len(c), len(r)).T
            pos += n
        assert pos == len(s)
        return out.T.reshape(-1)

    # per-spinor irrep labels in global order (for the generator)
    def occ_irreps(self):
        return np.repeat(np.arange(self.nrep), self.no)

    def vir_irreps(self):
        return np.repeat(np.arange(self.nrep), self.nv)
