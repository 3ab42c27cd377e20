"""CCSD Lambda equations and one-particle response density by exact
polynomial differentiation of the relccsd residual (test infrastructure only).

relccsd carries Lambda branches (CCEQN cceqns.F:277-288, T1EQNS :860-932,
the Lambda-only GETTAU2/3, LAMBD21, CINTM, DINTM in ccintm.F) but switches
them off (ccdriv.F:1918-1924) and calls them untested (cceqns.F:907), so they
specify the equations and pin nothing.  The equations are fixed by the CCSD
Lagrangian instead:

    L(T, Lam) = E(T) + sum_k Lam_k R_k(T),   R(T) = S(T) - D o T,

with S the right-hand side one CCEQN pass forms before DENOM (T1EQNS,
T2EQNS, BINTM and their intermediates), D the DENOM denominators
(cceqns.F:2397, :2435) and E the CCENG energy before its real part is taken
(cceqns.F:2180-2270), all in relccsd's packed coordinates (T1 then T2 over
A>B, I>J).  E and R are polynomials in T (no conjugate of T appears), so the
stationarity condition is holomorphic:

    dE/dT + Lam . dR/dT = 0   <=>   Lam = (dE/dT + Lam . dS/dT) / D,

and for a one-electron perturbation f -> f + lam V (V Hermitian; FOO, FVO,
FVV and the orbital energies shift by the blocks and the diagonal of V)

    d Ec / d lam = Re[ dE/dlam + Lam . dR/dlam ]  =  tr(G V)

with G the Hermitian one-particle response density.

Every derivative here is exact: a polynomial p of degree d in a scalar eps
has p'(0) = (1/N) sum_k w^-k p(w^k), w = exp(2 pi i / N), for any N >= d
(the discrete Fourier coefficient of order one, which aliases only orders
1 + N, 1 + 2N, ...; no truncation error).  R has
degree 4 in T (tau.A carries T1^4), E degree 2; both are affine in the Fock
blocks.  The dense Jacobian this gives limits the module to small systems.

Pinned by tests/test_oracle_lambda.py: central finite differences of the
oracle's converged CCSD energy, and of the relccsd reference's converged
energy (tests/golden/lambda_*.npz), along seeded Hermitian V.
"""
from __future__ import annotations

import numpy as np

from . import ccsd_oracle as co

NPTS = 6      # >= degree of the residual in T (4)


def rhs(S, t1, t2):
    """One CCEQN pass before DENOM (cceqns.F:313-483): dense (s1, s2)."""
    ho = co.hointm(S, t1, t2)
    hv = co.hvintm(S, t1, t2)
    hov = co.hovintm(S, t1)
    s1 = co.t1eqns(S, t1, t2, ho, hv, hov)
    go = co.gointm(S, t1, ho)
    gv = co.gvintm(S, t1, hv)
    az = co.aintm(S, t1, t2)
    h = co.hintm(S, t1, t2)
    s2 = co.t2eqns(S, t1, t2, go, gv, az, h)
    return s1, co.bintm(S, t1, t2, s2)


class Problem:
    """The residual and energy of one system as functions of the packed
    amplitude vector x = [T1 | T2]."""

    def __init__(self, sysd):
        self.S = co.System(sysd)
        L = self.L = self.S.L
        self.nt1 = L.nvo
        self.nt2 = L.nvvt * L.noot
        eo, ev = self.S.eps[:L.no], self.S.eps[L.no:]
        d1 = eo[None, :] - ev[:, None]
        d2 = (eo[None, None, :, None] + eo[None, None, None, :]
              - ev[:, None, None, None] - ev[None, :, None, None])
        self.D = np.concatenate([L.pt1(d1), L.pt2(d2)])
        self.fvo = L.pt1(self.S.fvo)
        self.vvoo = L.pt2(self.S.vvoo)

    def split(self, x):
        return self.L.t1(x[:self.nt1]), self.L.t2(x[self.nt1:])

    def residual(self, x):
        t1, t2 = self.split(x)
        s1, s2 = rhs(self.S, t1, t2)
        return np.concatenate([self.L.pt1(s1), self.L.pt2(s2)]) - self.D * x

    def energy(self, x):
        """CCENG's complex sum zdotc(FVO,T1) + zdotc(VVOO,tau)."""
        t1, t2 = self.split(x)
        return np.vdot(self.fvo, x[:self.nt1]) + np.vdot(self.vvoo, self.L.pt2(co.gettau(t1, t2)))


def first_coefficient(f, x, u, npts=NPTS):
    """d/deps f(x + eps u) at eps = 0 for a polynomial f of degree <= npts."""
    w = np.exp(2j * np.pi * np.arange(npts) / npts)
    acc = 0
    for k in range(npts):
        acc = acc + np.conj(w[k]) * f(x + w[k] * u)
    return acc / npts


def jacobian(P, x, npts=NPTS):
    """dR/dx (holomorphic), dense n x n."""
    n = x.size
    J = np.empty((n, n), dtype=np.complex128)
    e = np.zeros(n, dtype=np.complex128)
    for j in range(n):
        e[j] = 1.0
        J[:, j] = first_coefficient(P.residual, x, e, npts)
        e[j] = 0.0
    return J


def energy_gradient(P, x):
    n = x.size
    g = np.empty(n, dtype=np.complex128)
    e = np.zeros(n, dtype=np.complex128)
    for j in range(n):
        e[j] = 1.0
        g[j] = first_coefficient(P.energy, x, e, 3)
        e[j] = 0.0
    return g


def solve_lambda(sysd, T1, T2):
    """Lam (packed [L1 | L2]) at the amplitudes T: Lam J = -dE/dT."""
    P = Problem(sysd)
    x = np.concatenate([np.asarray(T1, np.complex128), np.asarray(T2, np.complex128)])
    J = jacobian(P, x)
    g = energy_gradient(P, x)
    lam = np.linalg.solve(J.T, -g)
    return lam[:P.nt1], lam[P.nt1:]


# --------------------------------------------------------------------------
# perturbations and the density


def _blocks(no, V):
    o, v = slice(0, no), slice(no, V.shape[0])
    return V[o, o], V[v, o], V[v, v]


def perturbed(sysd, V, lam):
    """The system with f -> f + lam V: FOO/FVO/FVV (column-major blocks of f)
    and the orbital energies (the diagonal of f) shifted."""
    no = int(sysd["no"])
    out = dict(sysd)
    voo, vvo, vvv = _blocks(no, V)
    for key, blk in (("FOO", voo), ("FVO", vvo), ("FVV", vvv)):
        out[key] = co._np(sysd[key]).astype(np.complex128) + lam * blk.T.reshape(-1)
    out["eps"] = co._np(sysd["eps"]).astype(np.float64) + lam * np.real(np.diag(V))
    return out


def fock_only(sysd, V):
    """A system whose only Hamiltonian part is the one-electron V (integrals
    zero); R and E are affine in the Fock blocks, so its residual and energy
    are dR/dlam and dE/dlam."""
    out = perturbed({k: (np.zeros_like(co._np(v)) if k in ("FOO", "FVO", "FVV", "eps") else v)
                     for k, v in sysd.items()}, V, 1.0)
    for c in ("OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV"):
        out[c] = np.zeros_like(co._np(sysd[c]))
    return out


def directional_derivative(sysd, T1, T2, L1, L2, V):
    """Re[dE/dlam + Lam . dR/dlam] at (T, Lam) along V."""
    P = Problem(fock_only(sysd, V))
    x = np.concatenate([np.asarray(T1, np.complex128), np.asarray(T2, np.complex128)])
    lam = np.concatenate([np.asarray(L1, np.complex128), np.asarray(L2, np.complex128)])
    return float(np.real(P.energy(x) + lam @ P.residual(x)))


def density(sysd, T1, T2, L1, L2):
    """The Hermitian response density G (n x n, spinors occupied then
    virtual) with d Ec / d lam = tr(G V) for every Hermitian V, probed on the
    Hermitian basis e_pp, e_pq + e_qp, i e_pq - i e_qp."""
    no, nv = int(sysd["no"]), int(sysd["nv"])
    n = no + nv
    G = np.zeros((n, n), dtype=np.complex128)
    for p in range(n):
        V = np.zeros((n, n), dtype=np.complex128)
        V[p, p] = 1.0
        G[p, p] = directional_derivative(sysd, T1, T2, L1, L2, V)
        for q in range(p + 1, n):
            V = np.zeros((n, n), dtype=np.complex128)
            V[p, q] = V[q, p] = 1.0
            re = directional_derivative(sysd, T1, T2, L1, L2, V)
            V[p, q], V[q, p] = 1j, -1j
            im = directional_derivative(sysd, T1, T2, L1, L2, V)
            G[p, q] = 0.5 * (re + 1j * im)
            G[q, p] = np.conj(G[p, q])
    return G


def hermitian_direction(n, seed, sigma=0.01):
    """The seeded Hermitian perturbation of the synthetic-input generator."""
    from molfdir_amd.synth import hermitian_perturbation
    return hermitian_perturbation(n, seed, sigma)


def fd_gradient(energy_of, sysd, V, h=1e-4):
    """Central difference (E(+h) - E(-h)) / 2h of a converged energy."""
    return (energy_of(perturbed(sysd, V, h)) - energy_of(perturbed(sysd, V, -h))) / (2 * h)
