"""relccsd's DIIS extrapolation restated on the CPU (test infrastructure).

Follows DIISX (cceqns.F:1985-2178) and FLIN/DOTT (ccgenp.F:666-762)
operation by operation:

* error d = S_new - T_old over T1 and T2 jointly (cceqns.F:2028-2031);
* RMS = sqrt(| sum|d|^2 - |sum d|^2 |), not normalised (cceqns.F:2034-2064);
* circular store of MAXDIM (error, new-amplitude) pairs, slot ITC cycling
  1..MAXDIM, subspace size IT = min(iteration, MAXDIM) (cceqns.F:2022-2026);
* B(n, ITC) = B(ITC, n) = Re<d_n|d_ITC> + Im<d_n|d_ITC> (cceqns.F:2078-2103);
* B scaled by 1/max|B| over its lower triangle, bordered with -1, solved by
  FLIN (Crout elimination with row pivoting, ccgenp.F:666-740);
* S = sum_n c_n S_n over the stored new amplitudes (cceqns.F:2155-2176).
"""
from __future__ import annotations

import numpy as np


class DiisState:
    def __init__(self, maxdim):
        self.maxdim = maxdim
        self.it = 0
        self.itc = 0
        self.err = [None] * maxdim      # slot -> (d1, d2)
        self.amp = [None] * maxdim      # slot -> (s1, s2)
        self.bb2 = np.zeros((maxdim + 1, maxdim + 2))
        self.last_coef = None


def _dott(a, b):
    d = 0.0
    for x, y in zip(a, b):
        d = d + x * y
    return d


def flin(a, n, nr):
    """FLIN (ccgenp.F:666-740) on a (IDIM x n+nr) array, in place.
    The solution overwrites columns n..n+nr-1.  Returns det."""
    jmax = n + nr
    sign = 1.0
    d = 0.0
    for m in range(n):
        temp = 0.0
        imax = m
        for i in range(m, n):
            if m > 0:
                a[i, m] = a[i, m] - _dott(a[i, :m], a[:m, m])
            aval = a[i, m]
            if abs(aval) <= temp:
                continue
            temp = abs(aval)
            imax = i
        if temp <= 0.0:
            return 0.0
        if imax != m:
            sign = -sign
            a[[m, imax], :jmax] = a[[imax, m], :jmax]
        jj = m + 1
        if jj >= jmax:
            continue
        if m == 0:
            a[0, jj:jmax] = a[0, jj:jmax] / a[0, 0]
            d = a[0, 0]
            continue
        for j in range(jj, jmax):
            a[m, j] = (a[m, j] - _dott(a[m, :m], a[:m, j])) / a[m, m]
        d = d * a[m, m]
    for i in range(nr):
        npi = n + i
        for k in range(2, n + 1):
            j = n - k
            a[j, npi] = a[j, npi] - _dott(a[j, j + 1:n], a[j + 1:n, npi])
    return d * sign


def rms_of(d1, d2):
    csum = d1.sum() + d2.sum()
    sum2 = np.real(np.vdot(d1, d1)) + np.real(np.vdot(d2, d2))
    return float(np.sqrt(abs(sum2 - np.real(np.conj(csum) * csum))))


def solve_coefficients(bb2, it):
    """Scale, border and solve the DIIS system as DIISX does; returns c."""
    bb = np.zeros((it + 1, it + 2))
    bb[:it, :it] = bb2[:it, :it]
    xm = abs(bb[0, 0])
    for n in range(it):
        for m in range(n + 1):
            xm = max(xm, abs(bb[n, m]))
    bb[:it, :it] *= 1.0 / xm
    bb[:it, it] = -1.0
    bb[it, :it] = -1.0
    bb[it, it] = 0.0
    bb[:it, it + 1] = 0.0
    bb[it, it + 1] = -1.0
    flin(bb, it + 1, 1)
    return bb[:it, it + 1].copy()


def diisx(st: DiisState, t1o, t1n, t2o, t2n):
    """One DIISX call.  Returns (S1, S2, rms) with S the extrapolated
    amplitudes.  Inputs are packed complex vectors (not modified)."""
    st.it = min(st.it + 1, st.maxdim)
    st.itc = st.itc + 1
    if st.itc > st.maxdim:
        st.itc = 1
    d1 = t1n - t1o
    d2 = t2n - t2o
    rms = rms_of(d1, d2)
    c = st.itc - 1
    st.err[c] = (d1.copy(), d2.copy())
    st.amp[c] = (t1n.copy(), t2n.copy())
    for n in range(st.it):
        e1, e2 = st.err[n]
        z = np.vdot(e1, d1) + np.vdot(e2, d2)
        v = float(np.real(z) + np.imag(z))
        st.bb2[n, c] = v
        st.bb2[c, n] = v
    coef = solve_coefficients(st.bb2, st.it)
    st.last_coef = coef
    s1 = np.zeros_like(t1n)
    s2 = np.zeros_like(t2n)
    for n in range(st.it):
        a1, a2 = st.amp[n]
        s1 = s1 + coef[n] * a1
        s2 = s2 + coef[n] * a2
    return s1, s2, rms
