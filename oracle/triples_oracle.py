"""CPU restatement of relccsd's non-iterative triples corrections T3CORR
(cceqns.F:1336-1863) for the C1 double group (test infrastructure only).

T3CORR follows Deegan and Knowles, CPL 227 (1994) 321.  For every occupied
triple i > j > k and virtual triple a > b > c it forms

  X(ij,k)(a;bc) = sum_e T(ae;ij) conj<ek||bc> + sum_l <al||ij> T(bc;kl)
  U'(a;bc)      = X(ij,k) + X(jk,i) - X(ik,j)                 (6 GEMMs, 1471-1561)
  U(abc)        = U'(a;bc) - U'(b;ac) + U'(c;ab)              (1580-1614)
  Y2'(ab;c)     = V(ab;ij) T1(c,k) + T(ab;ij) F(c,k)
                + V(ab;jk) T1(c,i) + T(ab;jk) F(c,i)
                - V(ab;ik) T1(c,j) - T(ab;ik) F(c,j)          (1615-1644)
  Y2(abc)       = Y2'(ab;c) + Y2'(bc;a) - Y2'(ac;b)           (1645-1680)
  Y3            = the same with TAU3 and T1 only, where
  TAU3(ab;ij)   = T(ab;ij) + (T1(a,i) T1(b,j) - T1(a,j) T1(b,i)) / 3
                                                    (GETTAU3, ccintm.F:1055-1153)
  D             = eps_a + eps_b + eps_c - eps_i - eps_j - eps_k

and accumulates (1681-1789), real parts of complex products:

  ET1 = -sum |U|^2 / D           fourth-order +T
  ET2 = -sum Re(conj(U) Y2) / D  fifth-order (T) term
  ET3 = +sum Re(conj(U) Y3)      fifth-order -T term

t3corr() is the restricted-loop restatement (one (a, bc) block per occupied
triple, like the reference); t3corr_full() is an independent dense
six-index form for tiny systems: the connected and disconnected triples
built with the antisymmetrisers P(k/ij) P(a/bc) over all index orders and
summed without restriction (1/36 of the full sum).

Pinned by tests/test_oracle_triples.py: against the compiled reference's
own T3CORR on seeded systems and amplitudes (tests/golden/t3_*.npz, C1 and
symmetry-blocked), and t3corr against t3corr_full.
"""
from __future__ import annotations

import numpy as np

from .ccsd_oracle import Layout, _np


class _Blocks:
    """The dense blocks T3CORR reads (System without OOOO, VOVO, VVVV)."""

    def __init__(self, sysd):
        no, nv = int(sysd["no"]), int(sysd["nv"])
        L = self.L = Layout(no, nv)
        self.no, self.nv = no, nv
        self.eps = _np(sysd["eps"]).astype(np.float64)
        self.fvo = L.mat(sysd["FVO"], nv, no)
        self.vooo = L.unpack_vo_pp(sysd["VOOO"], no)             # <ak||ij>
        self.vvoo = L.unpack_pp(sysd["VVOO"], nv, no)            # <ab||ij>
        self.vovv = L.unpack_vo_pp(sysd["VOVV"], nv)             # <ak||cd>


def _tau3(t1, t2):
    x = np.einsum("ai,bj->abij", t1, t1) / 3.0
    return t2 + x - x.transpose(0, 1, 3, 2)


def t3corr(sysd, T1, T2):
    """(ET1, ET2, ET3) of T3CORR for packed C1 amplitudes T1, T2."""
    S = _Blocks(sysd)
    L = S.L
    no, nv = S.no, S.nv
    t1, t2 = L.t1(T1), L.t2(T2)
    tau3 = _tau3(t1, t2)
    W = np.conj(S.vovv)                     # W[e,k,b,c] = conj <ek||bc>
    eo, ev = S.eps[:no], S.eps[no:]
    a_, b_, c_ = np.nonzero(np.ones((nv, nv, nv), dtype=bool))
    sel = (a_ > b_) & (b_ > c_)
    a_, b_, c_ = a_[sel], b_[sel], c_[sel]
    dv = ev[a_] + ev[b_] + ev[c_]
    e1 = e2 = e3 = 0.0
    for k in range(no):
        for j in range(k + 1, no):
            for i in range(j + 1, no):

                def X(p, q, r):
                    a1 = np.ascontiguousarray(t2[:, :, p, q])
                    b1 = np.ascontiguousarray(W[:, r].reshape(nv, nv * nv))
                    a2 = np.ascontiguousarray(S.vooo[:, :, p, q])
                    b2 = np.ascontiguousarray(t2[:, :, r, :].reshape(nv * nv, no).T)
                    return (a1 @ b1 + a2 @ b2).reshape(nv, nv, nv)

                Up = X(i, j, k) + X(j, k, i) - X(i, k, j)          # U'[a,b,c], antisym in bc
                U = Up[a_, b_, c_] - Up[b_, a_, c_] + Up[c_, a_, b_]

                def Yp(G, F):
                    return (G[:, :, i, j, None] * F[None, None, :, k]
                            + G[:, :, j, k, None] * F[None, None, :, i]
                            - G[:, :, i, k, None] * F[None, None, :, j])

                Y2p = Yp(S.vvoo, t1) + Yp(t2, S.fvo)               # Y2'[a,b,c] = (ab;c)
                Y3p = Yp(tau3, t1)
                Y2 = Y2p[a_, b_, c_] + Y2p[b_, c_, a_] - Y2p[a_, c_, b_]
                Y3 = Y3p[a_, b_, c_] + Y3p[b_, c_, a_] - Y3p[a_, c_, b_]
                D = dv - eo[i] - eo[j] - eo[k]
                e1 -= np.sum(np.abs(U) ** 2 / D)
                e2 -= np.sum(np.real(np.conj(U) * Y2) / D)
                e3 += np.sum(np.real(np.conj(U) * Y3))
    return float(e1), float(e2), float(e3)


def _antisym_ijk(x):
    """P(k/ij) on the last three indices: f(ijk) - f(kji) - f(ikj)."""
    return x - x.transpose(0, 1, 2, 5, 4, 3) - x.transpose(0, 1, 2, 3, 5, 4)


def _antisym_abc(x):
    """P(a/bc) on the first three indices: f(abc) - f(bac) - f(cba)."""
    return x - x.transpose(1, 0, 2, 3, 4, 5) - x.transpose(2, 1, 0, 3, 4, 5)


def t3corr_full(sysd, T1, T2):
    """Dense six-index form (tiny systems only): W[a,b,c,i,j,k] connected,
    Z2 / Z3 disconnected, all fully antisymmetric; each correction is 1/36
    of the unrestricted sum."""
    S = _Blocks(sysd)
    L = S.L
    no, nv = S.no, S.nv
    t1, t2 = L.t1(T1), L.t2(T2)
    tau3 = _tau3(t1, t2)
    # connected: sum_e t(ae;ij) <bc||ek> + sum_l <al||ij> t(bc;kl), with the
    # odd occupied index k and the odd virtual a, antisymmetrised over both
    x = (np.einsum("aeij,ekbc->abcijk", t2, np.conj(S.vovv))
         + np.einsum("alij,bckl->abcijk", S.vooo, t2))
    W = _antisym_abc(_antisym_ijk(x))
    # disconnected: (ab;ij) times (c;k), odd indices c and k
    def z(G, F):
        y = np.einsum("abij,ck->abcijk", G, F)
        y = y - y.transpose(0, 1, 2, 5, 4, 3) - y.transpose(0, 1, 2, 3, 5, 4)
        return y - y.transpose(2, 1, 0, 3, 4, 5) - y.transpose(0, 2, 1, 3, 4, 5)
    Z2 = z(S.vvoo, t1) + z(t2, S.fvo)
    Z3 = z(tau3, t1)
    e = S.eps
    eo, ev = e[:no], e[no:]
    D = (ev[:, None, None, None, None, None] + ev[None, :, None, None, None, None]
         + ev[None, None, :, None, None, None] - eo[None, None, None, :, None, None]
         - eo[None, None, None, None, :, None] - eo[None, None, None, None, None, :])
    Dinv = 1.0 / D
    e1 = -np.sum(np.abs(W) ** 2 * Dinv) / 36.0
    e2 = -np.sum(np.real(np.conj(W) * Z2) * Dinv) / 36.0
    e3 = np.sum(np.real(np.conj(W) * Z3)) / 36.0
    return float(e1), float(e2), float(e3)
