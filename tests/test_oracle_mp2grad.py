"""The MP2 relaxed property oracle (oracle/mp2grad_oracle.py) against
relccsd's own chain: MP2EQN 'LAMBDA', CCDENSC(1), CCDENSZ(1) with the RPAA
Z-vector and EXPVAL, run by harness mode 7 on seeded systems
(tests/golden/mp2g_*.npz: 4o/10v and 6o/20v with a non-canonical Fock
matrix, 8o/30v canonical; 6o/20v with 2 irreps and 7o/24v with 4 irreps and
a totally symmetric operator).

* MP2 amplitudes, DOO, DVV, the X intermediate and the preconditioner DG
  agree to rounding (1e-13 relative).
* The Z-vector agrees to the reference's convergence (RPAA stops at a DIIS
  RMS below 1e-13): 1e-10 relative; the expectation value to 1e-12.
* A planted mistake (the VOVV x DVV term of X left out, or B of the RPA
  matrix with the wrong sign) is caught.
"""
import numpy as np
import pytest

from molfdir_amd import synth
from oracle import mp2grad_oracle as O
from oracle.ccsd_oracle import Layout, System

from _golden import layout_for, load, names, system_for


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300)


def case(name):
    g = load("mp2g", name)
    no, nv, _, pseed = (int(x) for x in g["params"][:4])
    s = system_for(g)
    irreps = synth.spinor_irreps(s["sym"]) if "sym" in s else None
    P = synth.hermitian_perturbation(no + nv, pseed, 0.1, irreps=irreps)
    return g, s, P, Layout(no, nv)


def c1_names():
    return [n for n in names("mp2g") if "sym_no" not in load("mp2g", n)]


@pytest.mark.parametrize("name", names("mp2g"))
def test_chain_matches_reference(name):
    """C1 fixtures directly; symmetric ones (2 and 4 irreps) through SYMTAB's
    block maps, the VO vectors over the totally symmetric block (the other
    NDIMX pairs of the reference's vectors are zero)."""
    g, s, P, L = case(name)
    o = O.mp2_relaxed(s, P)
    conv = {"T1": L.pt1(o["t1"]), "DOO": L.pmat(o["doo"]), "DVV": L.pmat(o["dvv"]),
            "XVO": o["xvo"], "DVO": o["dvo"], "DOV": o["dov"]}
    nt1 = conv["T1"].size
    if "sym_no" in g:
        lay = layout_for(g)
        conv = {k: lay.to_sym({"DOO": "FOO", "DVV": "FVV"}.get(k, "T1"), v)
                for k, v in conv.items()}
        nt1 = lay.size("T1")
        for k in ("XVO", "DVO", "DOV"):
            assert np.max(np.abs(g[k][nt1:]), initial=0.0) < 1e-14
    assert rel(conv["T1"], g["T1"]) < 1e-13
    assert rel(conv["DOO"], g["DOO"]) < 1e-13
    assert rel(conv["DVV"], g["DVV"]) < 1e-13
    assert rel(conv["XVO"], g["XVO"][:nt1]) < 1e-13
    assert rel(conv["DVO"], g["DVO"][:nt1]) < 1e-10
    assert rel(conv["DOV"], g["DOV"][:nt1]) < 1e-10
    if "sym_no" not in g:
        assert rel(o["dg"], g["DG"]) < 1e-14
    assert abs(o["expval"] - complex(g["expval"])) < 1e-12


@pytest.mark.parametrize("name", c1_names())
def test_zvector_structure(name):
    """Y = conj X, and the reference's EXPVAL equals the oracle's formula on
    the reference's own densities."""
    g, s, P, L = case(name)
    assert rel(g["DOV"], np.conj(g["DVO"])) < 1e-12
    no = int(g["params"][0])
    nv = P.shape[0] - no
    e = O.expval(P, no, g["DOO"].reshape(no, no).T, g["DVV"].reshape(nv, nv).T, g["DVO"],
                 g["DOV"])
    assert abs(e - complex(g["expval"])) < 1e-14


def test_planted_mistakes_are_caught():
    g, s, P, L = case("g_6o20v")
    S = System(s)
    t1, t2, l1, l2 = O.mp2_lambda(S)
    doo, dvv = O.ccdensc(t1, t2, l1, l2)
    good = -O.vo(O.xintm(S, l2, doo, dvv))
    bad = -O.vo(O.xintm(S, l2, doo, np.zeros_like(dvv)))
    assert rel(good, g["XVO"]) < 1e-13
    assert rel(bad, g["XVO"]) > 1e-3
    m, bb, _ = O.rpa_matrices(S)
    n = m.shape[0]
    big = np.block([[m, -bb], [-np.conj(bb), np.conj(m)]])
    sol = np.linalg.solve(big, np.concatenate([good, np.conj(good)]))
    assert rel(sol[:n], g["DVO"]) > 1e-3
