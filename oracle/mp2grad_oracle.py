"""CPU restatement of relccsd's MP2 relaxed first-order property (test
infrastructure only): the DOMP2G block of CCFOPR (ccdriv.F:2058-2090) for
the C1 double group, in dense numpy tensors.

* MP2EQN with EQNS = 'LAMBDA' (cceqns.F:33-139): T1 = FVO / D, T2 = VVOO / D,
  and the lambda amplitudes are a copy of them (cceqns.F:116-117).
* CCDENSC(1) (ccgrad.F:34-100): the amplitude part of the one-particle
  density,
      DOO(i,j) = conj[ -1/2 sum_cdl conj T(cd,il) L(cd,jl) - sum_c conj T(c,i) L(c,j) ]
      DVV(a,b) = conj[  1/2 sum_dkl T(ad,kl) conj L(bd,kl) + sum_k T(a,k) conj L(b,k) ].
* CCDENSZ(1) (ccgrad.F:141-259): the X intermediate
      X(a,i) = 1/2 sum <kl||ci> L(ac,kl)              (T1EQN1 with tau -> L)
             + 1/2 sum <ak||cd> L(cd,ik)              (T1EQN2 with tau -> L)
             + sum_kl <ak||il> DOO(k,l)               (SRT1ST4 + XGEMV)
             + sum_bc DVV(b,c) conj <bi||ca>          (SRT20D, left-hand)
  stored negated (XSCAL -1), and XOV = conj(XVO).
* RPAA (ccrpa.F:34-282) at zero frequency, with DOAX / DOBY / GETDG / PRECD
  (ccrpa.F:467-603, 695-754, 1450-1528): its fixed point is
      M X + B Y = XVO,  conj(B) X + conj(M) Y = XOV,
      M(ai,ck) = (eps_a - eps_i) delta - <ak||ci>,  B(ai,ck) = <ac||ik>,
  solved here directly (RPAA iterates it with DIIS to RMS < 10^-NTOLGR);
  DVO = X, DOV = Y, and DG(ai) = eps_a - eps_i - <ai||ai>.
* EXPVAL (ccgrad.F:261-347): sum P(i,a) DVO(a,i) + P(a,i) DOV(a,i)
  + P(i,j) DOO(i,j) + P(a,b) DVV(a,b) + sum_i P(i,i) (unconjugated dot
  products of RDPROP's blocks, EPHASE = 1 for these operators).

Pinned by tests/test_oracle_mp2grad.py against the reference's own chain
(harness mode 7, tests/golden/mp2g_*.npz): every intermediate and the value.
"""
from __future__ import annotations

import numpy as np

from .ccsd_oracle import System, _es, denom


def mp2_lambda(S: System):
    """MP2EQN 'LAMBDA': amplitudes and lambda (a copy) as dense t1[a,i], t2."""
    t1, t2 = denom(S, S.fvo, S.vvoo)
    return t1, t2, t1.copy(), t2.copy()


def ccdensc(t1, t2, l1, l2):
    """CCDENSC(1): doo[i,j], dvv[a,b]."""
    doo = -0.5 * _es("cdil,cdjl->ij", np.conj(t2), l2) - np.conj(t1).T @ l1
    dvv = 0.5 * _es("adkl,bdkl->ab", t2, np.conj(l2)) + t1 @ np.conj(l1).T
    return np.conj(doo), np.conj(dvv)


def xintm(S: System, l2, doo, dvv):
    """CCDENSZ's X intermediate x[a,i] before the sign change."""
    klci = np.conj(S.vooo.transpose(2, 3, 0, 1))          # <kl||ci>
    x = 0.5 * _es("klci,ackl->ai", klci, l2)
    x += 0.5 * _es("akcd,cdik->ai", S.vovv, l2)
    x += _es("akil,kl->ai", S.vooo, doo)
    x += _es("bc,bica->ai", dvv, np.conj(S.vovv))
    return x


def rpa_matrices(S: System):
    """M and B over VO pairs (index i*nv + a) and the preconditioner DG."""
    no, nv = S.no, S.nv
    n = no * nv
    eo, ev = S.eps[:no], S.eps[no:]
    aa = S.vovo.transpose(0, 3, 2, 1)                       # [a,i,c,k] = <ak||ci>
    aa = aa.transpose(1, 0, 3, 2).reshape(n, n)             # rows (i,a), cols (k,c)
    bb = S.vvoo.transpose(0, 2, 1, 3)                       # [a,i,c,k] = <ac||ik>
    bb = bb.transpose(1, 0, 3, 2).reshape(n, n)
    de = (ev[None, :] - eo[:, None]).reshape(n)
    m = np.diag(de) - aa
    dg = de - np.real(np.diag(aa))
    return m, bb, dg


def zvector(S: System, xvo):
    """RPAA's converged (X, Y) for right-hand sides XVO, XOV = conj XVO
    (vectors over VO pairs, index i*nv + a)."""
    m, bb, _ = rpa_matrices(S)
    n = m.shape[0]
    big = np.block([[m, bb], [np.conj(bb), np.conj(m)]])
    sol = np.linalg.solve(big, np.concatenate([xvo, np.conj(xvo)]))
    return sol[:n], sol[n:]


def vo(x):
    """x[a,i] -> VO vector (index i*nv + a)."""
    return x.T.reshape(-1).copy()


def expval(P, no, doo, dvv, dvo, dov):
    """EXPVAL on dense doo[i,j], dvv[a,b] and VO vectors dvo, dov."""
    nv = P.shape[0] - no
    poo, pvv = P[:no, :no], P[no:, no:]
    pvo, pov = P[no:, :no], P[:no, no:]
    e = np.sum(vo(pov.T) * dvo) + np.sum(vo(pvo) * dov)
    e += np.sum(poo * doo) + np.sum(pvv * dvv) + np.trace(poo)
    assert dvv.shape == (nv, nv)
    return complex(e)


def mp2_relaxed(sysd, P):
    """The whole chain: dict of every intermediate (dense / VO vectors) and
    the expectation value of P (NSP x NSP, occupied first)."""
    S = System(sysd)
    t1, t2, l1, l2 = mp2_lambda(S)
    doo, dvv = ccdensc(t1, t2, l1, l2)
    x = -vo(xintm(S, l2, doo, dvv))
    dvo, dov = zvector(S, x)
    _, _, dg = rpa_matrices(S)
    return {"t1": t1, "t2": t2, "doo": doo, "dvv": dvv, "xvo": x, "dvo": dvo, "dov": dov,
            "dg": dg, "expval": expval(np.asarray(P), S.no, doo, dvv, dvo, dov)}
