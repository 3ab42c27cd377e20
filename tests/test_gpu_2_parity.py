"""GPU parity of the HIP path (through the C ABI) against the relccsd
reference fixtures and the CPU oracle.

Bars: per-term intermediates to 1e-12 relative; whole trajectories with the
correlation energy to 1e-10 Eh and amplitudes to 1e-9 max-abs (the FP64
parity tolerance of this path).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from molfdir_amd import synth  # noqa: E402
from molfdir_amd.relccsd import RelCCSD  # noqa: E402
from oracle import ccsd_oracle as O  # noqa: E402

from _golden import load, names, system_for, random_amplitudes  # noqa: E402

TERM_ORDER = [("hointm", ["HO"]), ("hvintm", ["HV"]), ("hovintm", ["HOV"]),
              ("t1eqns", ["S1"]), ("gointm", ["GO"]), ("gvintm", ["GV"]),
              ("aintm", ["AZ"]), ("hintm", ["H"]), ("t2eqns", ["S2"]), ("bintm", ["S2"])]
GOLDEN_KEY = {"HO": "HO", "HV": "HV", "HOV": "HOV", "S1": "S1", "GO": "GO", "GV": "GV",
              "AZ": "AZ", "H": "H"}


def solver_for(s, **kw):
    r = RelCCSD(int(s["no"]), int(s["nv"]), **kw)
    r.set_system(s)
    return r


@pytest.mark.parametrize("name", names("terms"))
def test_terms_match_reference(name):
    g = load("terms", name)
    no, nv, tseed = int(g["params"][0]), int(g["params"][1]), int(g["params"][3])
    s = system_for(g)
    t1, t2 = random_amplitudes(no, nv, tseed)
    r = solver_for(s)
    r.set_amplitudes(t1, t2)
    for term, outs in TERM_ORDER:
        r.term(term)
        for o in outs:
            if term == "bintm":
                key = "S2_BINTM"
            elif term == "t2eqns":
                key = "S2_T2EQNS"
            else:
                key = GOLDEN_KEY[o]
            got, ref = r.get(o), g[key]
            err = np.abs(got - ref).max()
            assert err <= 1e-12 * max(1.0, np.abs(ref).max()), (term, o, err)
    r.term("denom")
    r.term("zcore")
    np.testing.assert_allclose(r.get("S1"), g["S1_DENOM"], atol=1e-12, rtol=0)
    np.testing.assert_allclose(r.get("S2"), g["S2_DENOM"], atol=1e-12, rtol=0)
    assert abs(r.energy() - float(g["ECC"])) < 1e-12
    assert abs(r.t1diag() - float(g["T1DIAG"])) < 1e-13 * max(1.0, float(g["T1DIAG"]))


@pytest.mark.parametrize("three_mult", [True, False])
@pytest.mark.parametrize("name", names("ccsd"))
def test_trajectory_matches_reference(name, three_mult):
    g = load("ccsd", name)
    maxit, maxdim, ntol = (int(x) for x in g["params"][3:6])
    s = system_for(g)
    r = solver_for(s, maxdim=maxdim, maxit=maxit, ntol=ntol)
    r.set_complex_product(three_mult)
    out = r.cceqn()
    ref_rows = g["iterations"][1:]
    assert abs(out["emp2"] - float(g["emp2"])) < 1e-10
    assert out["nit"] == int(g["nit"]) == len(ref_rows)
    for (nit, e, rms), diag, ref in zip(out["rows"], out["t1diag"], ref_rows):
        assert abs(e - ref[1]) < 1e-10, (nit, e, ref[1])
        assert abs(rms - ref[2]) < 1e-9 * max(1.0, ref[2]), (nit, rms, ref[2])
        # T1-DIAGN column of the reference table (F10.5, cceqns.F:619)
        assert abs(diag - ref[3]) <= 5e-6 + 1e-12, (nit, diag, ref[3])
    t1, t2 = r.get_amplitudes()
    assert np.abs(t1 - g["T1"]).max() < 1e-9
    assert np.abs(t2 - g["T2"]).max() < 1e-9


@pytest.mark.parametrize("no,nv,seed", [(8, 52, 11), (10, 34, 12), (1, 9, 13), (7, 7, 14)])
def test_iterations_match_oracle_beyond_fixture_sizes(no, nv, seed):
    """Sizes and edge cases no fixture covers: single occupied (no T2),
    as many occupied as virtual, larger blocks that cross GEMM tiles."""
    s = synth.make_system(no, nv, seed=seed, scale=0.03, fock_perturb=0.02)
    ref = O.cceqn(s, maxit=6, maxdim=4, ntol=14)
    r = solver_for(s, maxdim=4, maxit=6, ntol=14)
    out = r.cceqn()
    assert abs(out["emp2"] - ref["emp2"]) < 1e-10
    for a, b in zip(out["rows"], ref["rows"]):
        assert abs(a[1] - b[1]) < 1e-10 and abs(a[2] - b[2]) < 1e-9 * max(1, b[2])
    for a, b in zip(out["t1diag"], ref["t1diag"]):
        assert abs(a - b) < 1e-12 * max(1.0, b)
    t1, t2 = r.get_amplitudes()
    assert np.abs(t1 - ref["T1"]).max() < 1e-9
    assert t2.size == ref["T2"].size
    if t2.size:
        assert np.abs(t2 - ref["T2"]).max() < 1e-9


def test_frozen_core_zeroes_amplitudes():
    no, nv = 5, 12
    s = synth.make_system(no, nv, seed=3, scale=0.03)
    r = solver_for(s, maxdim=4, maxit=3)
    r2 = RelCCSD(no, nv, frozen=[1, 0, 0, 0, 0], maxdim=4, maxit=3)
    r2.set_system(s)
    r2.cceqn()
    t1, t2 = r2.get_amplitudes()
    L = O.Layout(no, nv)
    assert np.all(t1.reshape(no, nv)[0] == 0)
    frozen_pairs = (L.jp == 0) | (L.ip == 0)
    assert np.all(t2.reshape(L.noot, L.nvvt)[frozen_pairs] == 0)
    assert np.any(t2.reshape(L.noot, L.nvvt)[~frozen_pairs] != 0)
    del r


def test_device_resident_inputs_match_host_inputs():
    no, nv = 6, 30
    s = synth.make_system(no, nv, seed=5, scale=0.03)
    sd = synth.make_system(no, nv, seed=5, scale=0.03, device="cuda")
    a = solver_for(s, maxdim=4, maxit=3)
    b = RelCCSD(no, nv, maxdim=4, maxit=3)
    b.set_fock(sd["FOO"], sd["FVO"], sd["FVV"], sd["eps"])
    for name in ("OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV"):
        b.set_ints(name, sd[name])
    ea, eb = a.cceqn(), b.cceqn()
    for x, y in zip(ea["rows"], eb["rows"]):
        assert abs(x[1] - y[1]) < 1e-12
