"""GPU parity of the frozen-core path (NFROZ != 0) and of T1DIAG against
relccsd runs with frozen occupied spinors.

relccsd marks the lowest NFROZ(irp) occupied spinors of each irrep as frozen
(IFROZ, SYMTAB ccdriv.F:838-849).  They stay in every intermediate, and
ZCORE (cceqns.F:1864-1922) zeroes T1(A,I) of frozen I and T2(AB,IJ) with I or
J frozen, on the MP2 guess (cceqns.F:94) and after every DENOM (:483).  The
handle takes the same mask (rcc_sym_t.frozen).  Bars as for every FP64
trajectory: Ec to 1e-10 Eh, RMS to 1e-9 relative, amplitudes to 1e-9; the
T1-diagnostic column to the 5 decimals relccsd prints, and to 1e-13 against
the full-precision value of the terms fixtures.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from molfdir_amd.relccsd import RelCCSD  # noqa: E402

from _golden import (load, names, system_for, layout_for, fixture_amplitudes,  # noqa: E402
                     frozen_mask_for, TERM_LAYOUT)

CLASSES = ("OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV")
TERM_ORDER = [("hointm", "HO"), ("hvintm", "HV"), ("hovintm", "HOV"), ("t1eqns", "S1"),
              ("gointm", "GO"), ("gvintm", "GV"), ("aintm", "AZ"), ("hintm", "H"),
              ("t2eqns", "S2_T2EQNS"), ("bintm", "S2_BINTM")]


def solver(g, **kw):
    """A handle on the fixture's system in the reference's layout (C1, or
    symmetry-blocked for 2-8 irreps), with the fixture's frozen mask."""
    s = system_for(g)
    no, nv = int(g["params"][0]), int(g["params"][1])
    fr = frozen_mask_for(g)
    if "sym_no" not in g:
        r = RelCCSD(no, nv, frozen=fr, **kw)
        r.set_system(s)
        return r, None
    lay = layout_for(g)
    r = RelCCSD(no, nv, frozen=fr, sym=(g["sym_no"].tolist(), g["sym_nv"].tolist()), **kw)
    r.set_fock(lay.to_sym("FOO", s["FOO"].numpy()), lay.to_sym("FVO", s["FVO"].numpy()),
               lay.to_sym("FVV", s["FVV"].numpy()), s["eps"])
    for name in CLASSES:
        r.set_ints(name, lay.to_sym(name, s[name].numpy()))
    return r, lay


@pytest.mark.parametrize("name", names("frzterms"))
def test_frozen_terms_match_reference(name):
    g = load("frzterms", name)
    r, lay = solver(g)
    t1, t2 = fixture_amplitudes(g)
    if lay is not None:
        t1, t2 = lay.to_sym("T1", t1), lay.to_sym("T2", t2)
    r.set_amplitudes(t1, t2)
    for term, key in TERM_ORDER:
        r.term(term)
        arr = {"S2_T2EQNS": "S2", "S2_BINTM": "S2"}.get(key, key)
        got, ref = r.get(arr), g[key]
        err = np.abs(got - ref).max()
        assert err <= 1e-12 * max(1.0, np.abs(ref).max()), (term, key, err)
    r.term("denom")
    r.term("zcore")
    np.testing.assert_allclose(r.get("S1"), g["S1_DENOM"], atol=1e-12, rtol=0)
    np.testing.assert_allclose(r.get("S2"), g["S2_DENOM"], atol=1e-12, rtol=0)
    # the frozen elements are exactly zero, as ZCORE's XCOPY of A0 leaves them
    assert np.array_equal(r.get("S1") == 0, g["S1_DENOM"] == 0)
    assert np.array_equal(r.get("S2") == 0, g["S2_DENOM"] == 0)
    assert abs(r.energy() - float(g["ECC"])) < 1e-12
    assert abs(r.t1diag() - float(g["T1DIAG"])) < 1e-13 * max(1.0, float(g["T1DIAG"]))


@pytest.mark.parametrize("three_mult", [True, False])
@pytest.mark.parametrize("name", names("frzccsd"))
def test_frozen_trajectory_matches_reference(name, three_mult):
    g = load("frzccsd", name)
    maxit, maxdim, ntol = (int(x) for x in g["params"][3:6])
    r, _ = solver(g, maxdim=maxdim, maxit=maxit, ntol=ntol)
    r.set_complex_product(three_mult)
    out = r.cceqn()
    ref_rows = g["iterations"][1:]
    assert abs(out["emp2"] - float(g["emp2"])) < 1e-10
    assert out["nit"] == int(g["nit"]) == len(ref_rows)
    for (nit, e, rms), diag, ref in zip(out["rows"], out["t1diag"], ref_rows):
        assert abs(e - ref[1]) < 1e-10, (nit, e, ref[1])
        assert abs(rms - ref[2]) < 1e-9 * max(1.0, ref[2]), (nit, rms, ref[2])
        assert abs(diag - ref[3]) <= 5e-6 + 1e-12, (nit, diag, ref[3])
    t1, t2 = r.get_amplitudes()
    assert t1.shape == g["T1"].shape and t2.shape == g["T2"].shape
    assert np.abs(t1 - g["T1"]).max() < 1e-9
    assert np.abs(t2 - g["T2"]).max() < 1e-9
    assert np.array_equal(t1 == 0, g["T1"] == 0)


@pytest.mark.parametrize("name", ["c1_5o16v", "s4_7o16v"])
def test_frozen_graph_replay_equals_eager(name):
    """The captured iteration keeps ZCORE: graph replays equal eager
    iterations bit for bit on a frozen-core handle."""
    g = load("frzccsd", name)
    maxit, maxdim, ntol = (int(x) for x in g["params"][3:6])
    a, _ = solver(g, maxdim=maxdim, maxit=maxit, ntol=ntol)
    b, _ = solver(g, maxdim=maxdim, maxit=maxit, ntol=ntol)
    b.set_graph(True)
    ea, eb = a.cceqn(), b.cceqn()
    assert ea["rows"] == eb["rows"]
    ta, tb = a.get_amplitudes(), b.get_amplitudes()
    assert np.array_equal(ta[0], tb[0]) and np.array_equal(ta[1], tb[1])
