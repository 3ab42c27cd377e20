"""The integral front-end oracle (oracle/rdints_oracle.py) against relccsd's
own sort of seeded MOLFDIR files (CCSETI -> RDINTS, tests/golden/sort_*.npz).

* The six classes agree bit for bit: the expansion moves values and flips
  signs only, and both sides keep the last write of an element in file order.
* The Fock blocks and orbital energies after GETFM / MAKEFM agree to 1e-14,
  E0 and E1 to the 15 decimals relccsd prints.
* A planted mistake -- keeping the first write of an element instead of the
  last -- fails on the lists with repeated images (the k5 list keeps every
  canonical integral, so many elements are written more than once).
"""
import hashlib
import os

import numpy as np
import pytest

from molfdir_amd import mdcint
from oracle import rdints_oracle as ro

from _golden import load, names

CLASSES = ("OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV")


def regenerate(g, d):
    nkr, nocc, seed = (int(x) for x in g["params"])
    gen = mdcint.synthetic_files(str(d), nkr, nocc, seed, float(g["keep"]))
    digest = hashlib.sha256(open(os.path.join(str(d), "MDCINT"), "rb").read()).hexdigest()
    assert digest == str(g["mdcint_sha256"]), "synthetic MDCINT generator changed"
    return gen, nkr, nocc


@pytest.mark.parametrize("name", names("sort"))
def test_oracle_sort_matches_reference(name, tmp_path):
    g = load("sort", name)
    gen, nkr, nocc = regenerate(g, tmp_path)
    sp = ro.Spinors(np.ones(2 * nkr, int), [nocc])
    kr = ro.kr_table(gen["kr"])
    cl = ro.sort_classes(sp, kr, gen["ikr"], gen["jkr"], gen["kkr"], gen["lkr"], gen["val"])
    for c in CLASSES:
        assert np.array_equal(cl[c], g[c]), c
    f = ro.fock(sp, gen["h1"], cl)
    for k in ("FOO", "FVO", "FVV"):
        assert np.abs(f[k] - g[k]).max() < 1e-14, k
    assert np.abs(f["eps"] - g["eps"]).max() < 1e-14
    assert abs(f["E0"] - float(g["E0"])) < 1e-13 and abs(f["E1"] - float(g["E1"])) < 1e-13
    assert abs(f["E0"] + f["E1"] - float(g["ESCF"])) < 1e-13


def test_planted_first_write_wins_is_caught(tmp_path, monkeypatch):
    g = load("sort", "k5_2o")
    gen, nkr, nocc = regenerate(g, tmp_path)
    sp = ro.Spinors(np.ones(2 * nkr, int), [nocc])
    kr = ro.kr_table(gen["kr"])
    seen = set()
    orig = ro.expand

    def first_wins(*a):
        out = []
        for e in orig(*a):
            key = (e[1] > 0, ro.CLASS_OF_TYPE.get(abs(e[1])), e[2], e[3])
            if key in seen:
                continue
            seen.add(key)
            out.append(e)
        return out

    monkeypatch.setattr(ro, "expand", first_wins)
    cl = ro.sort_classes(sp, kr, gen["ikr"], gen["jkr"], gen["kkr"], gen["lkr"], gen["val"])
    assert any(not np.array_equal(cl[c], g[c]) for c in CLASSES)


def regenerate_sym(g, d, breit=False):
    gen = mdcint.synthetic_files_sym(str(d), g["nkr_irrep"].tolist(), g["nocc_irrep"].tolist(),
                                     int(g["seed"]), float(g["keep"]), breit=breit)
    digest = hashlib.sha256(open(os.path.join(str(d), "MDCINT"), "rb").read()).hexdigest()
    assert digest == str(g["mdcint_sha256"]), "synthetic MDCINT generator changed"
    return gen


@pytest.mark.parametrize("name", names("symsort"))
def test_oracle_symmetric_sort_matches_reference(name, tmp_path):
    # classes in the C1 working order (irrep-major spinors), read out through
    # SYMTAB's block maps, against relccsd's symmetry-blocked sort
    from oracle.symtab import SymLayout
    g = load("symsort", name)
    gen = regenerate_sym(g, tmp_path)
    sp = ro.Spinors(gen["irreps"], gen["nelec"])
    kr = ro.kr_table(gen["kr"])
    cl = ro.sort_classes(sp, kr, gen["ikr"], gen["jkr"], gen["kkr"], gen["lkr"], gen["val"])
    lay = SymLayout(g["sym_no"].tolist(), g["sym_nv"].tolist())
    for c in CLASSES:
        assert np.array_equal(lay.to_sym(c, cl[c]), g[c]), c
    f = ro.fock(sp, gen["h1"], cl)
    for k in ("FOO", "FVO", "FVV"):
        assert np.abs(lay.to_sym(k, f[k]) - g[k]).max() < 1e-14, k
    assert np.abs(f["eps"] - g["eps"]).max() < 1e-14
    assert abs(f["E0"] - float(g["E0"])) < 1e-13 and abs(f["E1"] - float(g["E1"])) < 1e-13


@pytest.mark.parametrize("name", names("sortcont"))
def test_reference_reads_continuation_files(name, tmp_path):
    """The list of sort_<name> spread over MDCINT, MDCINT1, MDCINT2 (continuation
    records, GET2INTS ccints.F:832-838).  relccsd's first pass follows them:
    OOOO, VOOO, VVOO, VOVO, the Fock blocks and E0 / E1 equal the single-file
    sort bit for bit.  Its VOVV and VVVV passes restart with REWIND + one
    skipped record on the unit, which is then bound to the last file
    (ccints.F:291-292, 395-396), so those two classes hold exactly that file's
    records after its first one -- the library reads the whole list instead."""
    g = load("sortcont", name)
    single = load("sort", name)
    nkr, nocc, seed = (int(x) for x in g["params"])
    gen = mdcint.synthetic_files(str(tmp_path), nkr, nocc, seed, float(g["keep"]),
                                 parts=int(g["parts"]))
    digest = hashlib.sha256(open(os.path.join(str(tmp_path), "MDCINT"), "rb").read()).hexdigest()
    assert digest == str(g["mdcint_sha256"])
    for c in ("OOOO", "VOOO", "VVOO", "VOVO", "FOO", "FVO", "FVV", "eps"):
        assert np.array_equal(g[c], single[c]), c
    assert float(g["E0"]) == float(single["E0"]) and float(g["E1"]) == float(single["E1"])
    ik, jk = gen["ikr"], gen["jkr"]
    starts = [0] + [i for i in range(1, ik.size) if (ik[i], jk[i]) != (ik[i - 1], jk[i - 1])]
    parts = int(g["parts"])
    cuts = [round(i * len(starts) / parts) for i in range(parts + 1)]
    lo = starts[cuts[parts - 1] + 1]
    sel = slice(lo, ik.size)
    sp = ro.Spinors(np.ones(2 * nkr, int), [nocc])
    cl = ro.sort_classes(sp, ro.kr_table(gen["kr"]), ik[sel], jk[sel], gen["kkr"][sel],
                         gen["lkr"][sel], gen["val"][sel])
    for c in ("VOVV", "VVVV"):
        assert np.array_equal(cl[c], g[c]), c
        assert not np.array_equal(g[c], single[c]), c


@pytest.mark.parametrize("name", names("sortbreit"))
def test_oracle_breit_sort_matches_reference(name, tmp_path):
    """MRCONEE's BREIT flag set and a Breit list in MDBINT: relccsd sorts it
    with ITRSGN = -1 and adds the Coulomb classes (ccints.F:230-238, 496-505).
    The oracle agrees bit for bit; the Breit list sorted with ITRSGN = +1
    (a planted mistake) does not."""
    g = load("sortbreit", name)
    nkr, nocc, seed = (int(x) for x in g["params"])
    gen = mdcint.synthetic_files(str(tmp_path), nkr, nocc, seed, float(g["keep"]), breit=True)
    sp = ro.Spinors(np.ones(2 * nkr, int), [nocc])
    kr = ro.kr_table(gen["kr"])
    coul = (gen["ikr"], gen["jkr"], gen["kkr"], gen["lkr"], gen["val"])
    b = gen["breit"]
    brt = (b["ikr"], b["jkr"], b["kkr"], b["lkr"], b["val"])
    cl = ro.sort_classes_breit(sp, kr, coul, brt)
    for c in CLASSES:
        assert np.array_equal(cl[c], g[c]), c
    f = ro.fock(sp, gen["h1"], cl)
    assert abs(f["E0"] - float(g["E0"])) < 1e-13 and abs(f["E1"] - float(g["E1"])) < 1e-13
    wrong_b = ro.sort_classes(sp, kr, *brt)
    wrong_c = ro.sort_classes(sp, kr, *coul)
    assert any(not np.array_equal(wrong_b[c] + wrong_c[c], g[c]) for c in CLASSES)


@pytest.mark.parametrize("name", names("sortopt"))
def test_oracle_ccsort_options_match_reference(name, tmp_path):
    """Namelist CCSORT: USEOE keeps MRCONEE's orbital energies; NORECMP takes
    F = diag of them (FVO = 0), E0 = their occupied sum, E1 from the integrals
    (GETFM ccints.F:1080, 1150-1198, 1260, 1287).  The classes do not change."""
    g = load("sortopt", name)
    nkr, nocc, seed, useoe, norecmp = (int(x) for x in g["params"])
    gen = mdcint.synthetic_files(str(tmp_path), nkr, nocc, seed, float(g["keep"]))
    sp = ro.Spinors(np.ones(2 * nkr, int), [nocc])
    cl = ro.sort_classes(sp, ro.kr_table(gen["kr"]), gen["ikr"], gen["jkr"], gen["kkr"],
                         gen["lkr"], gen["val"])
    for c in CLASSES:
        assert np.array_equal(cl[c], g[c]), c
    f = ro.fock(sp, gen["h1"], cl, eps_file=gen["eps"], useoe=bool(useoe), norecmp=bool(norecmp))
    for k in ("FOO", "FVO", "FVV"):
        assert np.abs(f[k] - g[k]).max() < 1e-14, k
    assert np.abs(f["eps"] - g["eps"]).max() < 1e-14
    assert abs(f["E0"] - float(g["E0"])) < 1e-13 and abs(f["E1"] - float(g["E1"])) < 1e-13
    plain = ro.fock(sp, gen["h1"], cl)
    assert np.abs(plain["eps"] - g["eps"]).max() > 1e-6
