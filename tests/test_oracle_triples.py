"""The triples oracle (oracle/triples_oracle.py) against relccsd's own
T3CORR (cceqns.F:1336-1863) and against an independent dense form.

* tests/golden/t3_*.npz hold ET1, ET2, ET3 of the compiled reference on
  seeded systems: random amplitudes (C1, up to two 64-wide tiles of
  virtuals), the converged amplitudes of a reference CCSD run, symmetric
  systems with 4 and 8 fermion irreps (one irrep pair without occupied
  spinors), one virtual triple only (5o3v) and no occupied triple (2o6v).
* t3corr (restricted loops, one (a, bc) block per occupied triple, as the
  reference) against t3corr_full (six-index tensors antisymmetrised with
  P(k/ij) P(a/bc), unrestricted sums / 36): a wrong sign or index in the
  restricted bookkeeping of either shows up as a mismatch.
* A planted mistake (TAU3 without the 1/3, i.e. the ordinary tau) moves ET3
  outside the tolerance: the fixtures resolve that term.
"""
import numpy as np
import pytest

from oracle import triples_oracle as to

from _golden import load, names, random_amplitudes, symmetric_amplitudes, layout_for, system_for, \
    fingerprint


def amplitudes(g):
    no, nv = int(g["params"][0]), int(g["params"][1])
    amp = str(g["amplitudes"])
    if not amp.lstrip("-").isdigit():
        z = load("ccsd", amp)
        return z["T1"], z["T2"]
    if "sym_no" in g:
        return symmetric_amplitudes(layout_for(g), int(amp))
    return random_amplitudes(no, nv, int(amp))


def close(x, ref, rel=1e-12):
    return abs(x - ref) <= rel * max(1.0, abs(ref))


@pytest.mark.parametrize("name", names("t3"))
def test_oracle_matches_reference_t3corr(name):
    g = load("t3", name)
    s = system_for(g)
    assert np.allclose(fingerprint(s), g["fingerprint"], rtol=1e-12, atol=1e-12)
    t1, t2 = amplitudes(g)
    e = to.t3corr(s, t1, t2)
    for x, key in zip(e, ("ET1", "ET2", "ET3")):
        assert close(x, float(g[key])), (name, key, x, float(g[key]))


@pytest.mark.parametrize("name", ["r_4o10v", "r_5o3v", "s4_5o12v", "r_2o6v"])
def test_restricted_loops_equal_dense_six_index_form(name):
    g = load("t3", name)
    s = system_for(g)
    t1, t2 = amplitudes(g)
    a = to.t3corr(s, t1, t2)
    b = to.t3corr_full(s, t1, t2)
    for x, y in zip(a, b):
        assert close(x, y, 1e-13), (a, b)


def test_planted_tau3_mistake_is_caught(monkeypatch):
    g = load("t3", "r_4o10v")
    s = system_for(g)
    t1, t2 = amplitudes(g)

    def tau_without_third(t1, t2):
        x = np.einsum("ai,bj->abij", t1, t1)
        return t2 + x - x.transpose(0, 1, 3, 2)

    monkeypatch.setattr(to, "_tau3", tau_without_third)
    e = to.t3corr(s, t1, t2)
    assert close(e[0], float(g["ET1"])) and close(e[1], float(g["ET2"]))
    assert not close(e[2], float(g["ET3"]), 1e-6)
