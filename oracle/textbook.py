"""Independent dense spin-orbital CCSD (Stanton & Gauss, J. Chem. Phys. 94,
4334 (1991)) used to pin the oracle by mathematics, not by restating
relccsd.

Test infrastructure only.  It takes the full one-electron matrix f[p,q]
and antisymmetrised integrals g[p,q,r,s] = <pq||rs> built straight from the
seeded factor B (never from the packed classes), so a packing or
conjugation mistake in the layout code cannot cancel out.

Complex orbitals: every integral keeps its creation indices in the bra and
annihilation indices in the ket, so the published equations hold as written
except the driving terms, which for complex orbitals are <ab||ij> and f_ai
(not the real-orbital rewrites <ij||ab>, f_ia).

The Jacobi update divides by D = f_ii + f_jj - f_aa - f_bb with the orbital
energies eps (diagonal parts of F_ae / F_mi dropped), and the DIIS step is
the relccsd one (cceqns.F:1985-2178) so whole trajectories are comparable.
"""
from __future__ import annotations

import numpy as np

from . import diis as _diis


def dense_integrals(no, nv, seed, rank=120, scale=0.05):
    from molfdir_amd.synth import draw_factors
    eps, b = draw_factors(no, nv, seed, rank, scale)
    # (pr|qs) = sum_P B[P,p,r] B[P,q,s];  <pq|rs> = (pr|qs)
    mull = np.einsum("Ppr,Pqs->prqs", b, b, optimize=True)
    phys = mull.transpose(0, 2, 1, 3)              # <pq|rs>
    g = phys - phys.transpose(0, 1, 3, 2)          # <pq||rs>
    f = np.diag(eps).astype(np.complex128)
    return eps, f, g


def pack_t1(t1):
    """t1[a,i] -> relccsd T1 order (A fastest)."""
    return t1.T.reshape(-1).copy()


def pack_t2(t2):
    """t2[a,b,i,j] -> relccsd packed T2 (AB rows, IJ columns)."""
    nv, no = t2.shape[0], t2.shape[2]
    from molfdir_amd.synth import pair_index
    ap, bp = pair_index(nv)
    ip, jp = pair_index(no)
    return t2[ap[:, None], bp[:, None], ip[None, :], jp[None, :]].T.reshape(-1).copy()


def unpack_t1(v, no, nv):
    return v.reshape(no, nv).T.copy()


def unpack_t2(v, no, nv):
    from molfdir_amd.synth import pair_index
    ap, bp = pair_index(nv)
    ip, jp = pair_index(no)
    m = v.reshape(len(ip), len(ap)).T
    t = np.zeros((nv, nv, no, no), dtype=np.complex128)
    t[ap[:, None], bp[:, None], ip[None, :], jp[None, :]] = m
    t[bp[:, None], ap[:, None], ip[None, :], jp[None, :]] = -m
    t[ap[:, None], bp[:, None], jp[None, :], ip[None, :]] = -m
    t[bp[:, None], ap[:, None], jp[None, :], ip[None, :]] = m
    return t


class Textbook:
    def __init__(self, eps, f, g, no):
        self.no = no
        n = len(eps)
        self.nv = n - no
        o, v = slice(0, no), slice(no, n)
        self.eps = eps
        self.f = f
        self.g = g
        self.o, self.v = o, v
        eo, ev = eps[:no], eps[no:]
        self.D1 = (eo[None, :] - ev[:, None])                       # [a,i]
        self.D2 = (eo[None, None, :, None] + eo[None, None, None, :]
                   - ev[:, None, None, None] - ev[None, :, None, None])  # [a,b,i,j]

    def energy(self, t1, t2):
        o, v, f, g = self.o, self.v, self.f, self.g
        e = np.einsum("ia,ai->", f[o, v], t1)
        e += 0.25 * np.einsum("ijab,abij->", g[o, o, v, v], t2)
        e += 0.5 * np.einsum("ijab,ai,bj->", g[o, o, v, v], t1, t1)
        return e

    def update(self, t1, t2):
        """One Jacobi step: returns (t1_new, t2_new)."""
        o, v, f, g = self.o, self.v, self.f, self.g
        no, nv = self.no, self.nv
        fov = f[o, v]
        tau_t = t2 + 0.5 * (np.einsum("ai,bj->abij", t1, t1) - np.einsum("bi,aj->abij", t1, t1))
        tau = t2 + np.einsum("ai,bj->abij", t1, t1) - np.einsum("bi,aj->abij", t1, t1)
        Fae = f[v, v] * (1 - np.eye(nv))
        Fae = Fae - 0.5 * np.einsum("me,am->ae", fov, t1)
        Fae += np.einsum("fm,mafe->ae", t1, g[o, v, v, v])
        Fae -= 0.5 * np.einsum("afmn,mnef->ae", tau_t, g[o, o, v, v])
        Fmi = f[o, o] * (1 - np.eye(no))
        Fmi = Fmi + 0.5 * np.einsum("ei,me->mi", t1, fov)
        Fmi += np.einsum("en,mnie->mi", t1, g[o, o, o, v])
        Fmi += 0.5 * np.einsum("efin,mnef->mi", tau_t, g[o, o, v, v])
        Fme = fov + np.einsum("fn,mnef->me", t1, g[o, o, v, v])
        Wmnij = g[o, o, o, o].copy()
        x = np.einsum("ej,mnie->mnij", t1, g[o, o, o, v])
        Wmnij += x - x.transpose(0, 1, 3, 2)
        Wmnij += 0.25 * np.einsum("efij,mnef->mnij", tau, g[o, o, v, v])
        Wabef = g[v, v, v, v].copy()
        x = np.einsum("bm,amef->abef", t1, g[v, o, v, v])
        Wabef -= x - x.transpose(1, 0, 2, 3)
        Wabef += 0.25 * np.einsum("abmn,mnef->abef", tau, g[o, o, v, v])
        Wmbej = g[o, v, v, o].copy()
        Wmbej += np.einsum("fj,mbef->mbej", t1, g[o, v, v, v])
        Wmbej -= np.einsum("bn,mnej->mbej", t1, g[o, o, v, o])
        Wmbej -= np.einsum("fbjn,mnef->mbej",
                           0.5 * t2 + np.einsum("fj,bn->fbjn", t1, t1), g[o, o, v, v])

        # T1
        r1 = f[v, o].copy()
        r1 += np.einsum("ei,ae->ai", t1, Fae)
        r1 -= np.einsum("am,mi->ai", t1, Fmi)
        r1 += np.einsum("aeim,me->ai", t2, Fme)
        r1 -= np.einsum("fn,naif->ai", t1, g[o, v, o, v])
        r1 -= 0.5 * np.einsum("efim,maef->ai", t2, g[o, v, v, v])
        r1 -= 0.5 * np.einsum("aemn,nmei->ai", t2, g[o, o, v, o])
        # T2
        r2 = g[v, v, o, o].copy()
        x = np.einsum("aeij,be->abij", t2, Fae - 0.5 * np.einsum("bm,me->be", t1, Fme))
        r2 += x - x.transpose(1, 0, 2, 3)
        x = np.einsum("abim,mj->abij", t2, Fmi + 0.5 * np.einsum("ej,me->mj", t1, Fme))
        r2 -= x - x.transpose(0, 1, 3, 2)
        r2 += 0.5 * np.einsum("abmn,mnij->abij", tau, Wmnij)
        r2 += 0.5 * np.einsum("efij,abef->abij", tau, Wabef)
        x = np.einsum("aeim,mbej->abij", t2, Wmbej)
        x -= np.einsum("ei,am,mbej->abij", t1, t1, g[o, v, v, o])
        r2 += x - x.transpose(1, 0, 2, 3) - x.transpose(0, 1, 3, 2) + x.transpose(1, 0, 3, 2)
        x = np.einsum("ei,abej->abij", t1, g[v, v, v, o])
        r2 += x - x.transpose(0, 1, 3, 2)
        x = np.einsum("am,mbij->abij", t1, g[o, v, o, o])
        r2 -= x - x.transpose(1, 0, 2, 3)
        return r1 / self.D1, r2 / self.D2

    def mp2(self):
        o, v = self.o, self.v
        t1 = self.f[v, o] / self.D1
        t2 = self.g[v, v, o, o] / self.D2
        return t1, t2

    def run(self, maxit=20, maxdim=6, conv=1e-12):
        """relccsd's CCEQN loop (cceqns.F:293-569) on textbook residuals.
        Returns list of (nit, Ec, rms) and the final packed (T1, T2)."""
        no, nv = self.no, self.nv
        t1, t2 = self.mp2()
        T1, T2 = pack_t1(t1), pack_t2(t2)
        st = _diis.DiisState(maxdim)
        rows = []
        nit = 0
        while True:
            nit += 1
            if nit > maxit:
                break
            t1 = unpack_t1(T1, no, nv)
            t2 = unpack_t2(T2, no, nv)
            n1, n2 = self.update(t1, t2)
            S1, S2 = pack_t1(n1), pack_t2(n2)
            S1, S2, rms = _diis.diisx(st, T1, S1, T2, S2)
            ec = self.energy(unpack_t1(S1, no, nv), unpack_t2(S2, no, nv))
            rows.append((nit, ec, rms))
            T1, T2 = S1, S2
            if rms < conv:
                break
        return rows, T1, T2
