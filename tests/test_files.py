"""The MCCRES restart file (PUTAMPT / GETAMPT ccgetv.F:577-709, CCDUMP
ccdriv.F:2767-2816) through the library's host functions, against files
relccsd's CCENER wrote (tests/golden/mccres_*.bin, C1 and 4 irreps).

* Reading a reference file gives exactly the amplitudes, IMSTAT and energy
  records relccsd's own GETAMPT / CCDUMP records hold.
* Writing the same contents gives the same bytes wherever relccsd defines
  them (the status and energy prefixes, the T1 and T2 blocks); relccsd's
  GETAMPT was run on files of this writer when the fixtures were made.
* A file too short for the layout is refused.
* The library's MRCONEE writer reproduces the files relccsd reads in the
  integral-sort fixtures, byte for byte.
* The MDPROP writer's files are read by relccsd's RDPROP as the operator
  written (tests/golden/prop_*.npz).
"""
import os

import numpy as np
import pytest

from molfdir_amd.relccsd import RccError, mccres_read, mccres_write

from _golden import GOLDEN, load, names

REC = 8192


def regions(raw, nt1, nt2):
    nrec1 = nt1 * 2 // 1024 + 1
    t1 = raw[3 * REC:3 * REC + 16 * nt1]
    t2 = raw[(3 + nrec1) * REC:(3 + nrec1) * REC + 16 * nt2]
    return raw[REC:REC + 120], raw[2 * REC:2 * REC + 56], t1, t2


@pytest.mark.parametrize("name", names("mccres"))
def test_reader_gives_getampt_records(name):
    g = load("mccres", name)
    t1, t2, st, en = mccres_read(os.path.join(GOLDEN, f"mccres_{name}.bin"), g["T1"].size,
                                 g["T2"].size)
    assert np.array_equal(t1, g["T1"]) and np.array_equal(t2, g["T2"])
    assert np.array_equal(st, g["IMSTAT"]) and np.array_equal(en, g["energies"])
    assert st[4] == 3          # IMSTAT(5): CCSD done


@pytest.mark.parametrize("name", names("mccres"))
def test_writer_reproduces_reference_bytes(name, tmp_path):
    g = load("mccres", name)
    ref = open(os.path.join(GOLDEN, f"mccres_{name}.bin"), "rb").read()
    out = tmp_path / "MCCRES"
    mccres_write(out, g["T1"], g["T2"], imstat=g["IMSTAT"], energies=g["energies"])
    mine = out.read_bytes()
    assert len(mine) == len(ref)
    for a, b in zip(regions(mine, g["T1"].size, g["T2"].size),
                    regions(ref, g["T1"].size, g["T2"].size)):
        assert a == b


def test_short_file_is_refused(tmp_path):
    g = load("mccres", "s4_5o12v")
    p = tmp_path / "short"
    p.write_bytes(open(os.path.join(GOLDEN, "mccres_s4_5o12v.bin"), "rb").read()[:3 * REC])
    with pytest.raises(RccError):
        mccres_read(p, g["T1"].size, g["T2"].size)


@pytest.mark.parametrize("name", names("prop"))
def test_reference_rdprop_reads_the_mdprop_writer(name):
    """relccsd's RDPROP (harness mode 6) read 'PROPTEST' -- the second
    operator of an MDPROP file written by molfdir_amd.mdcint.write_mdprop --
    into AOO(i,j) = P(i,j), AVV(a,b) = P(a,b), AVO(a,i) = P(a,i) and
    AOV(a,i) = P(i,a), and EXPVAL on zero densities gave sum_i P(i,i)."""
    from molfdir_amd import synth
    g = load("prop", name)
    no, nv, _, pseed = (int(x) for x in g["params"])
    P = synth.hermitian_perturbation(no + nv, pseed, 0.1)
    assert np.array_equal(g["AOO"], P[:no, :no].T.reshape(-1))
    assert np.array_equal(g["AVV"], P[no:, no:].T.reshape(-1))
    assert np.array_equal(g["AVO"], P[no:, :no].T.reshape(-1))
    assert np.array_equal(g["AOV"], P[:no, no:].reshape(-1))
    assert abs(complex(g["expval"]) - np.trace(P[:no, :no])) < 1e-14


@pytest.mark.parametrize("nrep", [1, 2, 4])
def test_mrconee_writer_matches_the_files_the_reference_reads(nrep, tmp_path):
    """The library's MRCONEE writer gives, byte for byte, the files written by
    molfdir_amd.mdcint.write_mrconee, which relccsd's RDSYMI / GETFM read in
    the integral-sort fixtures (tests/golden/sort_*.npz, symsort_*.npz)."""
    from molfdir_amd import mdcint
    from molfdir_amd.relccsd import mrconee_write
    rng = np.random.default_rng(nrep)
    nsp = 12 * nrep
    irreps = rng.integers(1, nrep + 1, nsp).astype(np.int32)
    nelec = np.array([int(np.sum(irreps == r + 1)) // 3 for r in range(nrep)], np.int32)
    eps = rng.uniform(-2, 5, nsp)
    h1 = rng.normal(size=(nsp, nsp)) + 1j * rng.normal(size=(nsp, nsp))
    h1 = h1 + h1.conj().T
    multb = None if nrep == 1 else mdcint.abelian_multb(nrep)
    mdcint.write_mrconee(str(tmp_path / "py"), eps, irreps, nelec, h1, multb=multb, ecore=-3.25,
                         breit=nrep == 2)
    mrconee_write(tmp_path / "lib", eps, irreps, nelec, h1, multb=multb, ecore=-3.25,
                  breit=nrep == 2)
    assert (tmp_path / "lib").read_bytes() == (tmp_path / "py").read_bytes()
    with pytest.raises(RccError):
        mrconee_write(tmp_path / "bad", eps, irreps + nrep, nelec, h1, multb=multb)
