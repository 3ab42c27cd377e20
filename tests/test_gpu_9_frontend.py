"""GPU integral front end (rcc_read_molfdir, RDINTS ccints.F:33-569) against
relccsd's own sort of the same seeded MOLFDIR files (tests/golden/sort_*.npz).

Bar: the six classes bit for bit (values are moved and sign-flipped only;
the last write of an element in file order wins on both sides); Fock blocks
and orbital energies to 1e-14; E0 / E1 to 1e-12.  The device-list entry
point (rcc_sort_ints_device) gives the same classes.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from molfdir_amd.relccsd import RelCCSD, RccError  # noqa: E402

from _golden import load, names  # noqa: E402
from test_oracle_rdints import CLASSES, regenerate  # noqa: E402


@pytest.mark.parametrize("name", names("sort"))
def test_read_molfdir_matches_reference_sort(name, tmp_path):
    g = load("sort", name)
    print(f"sort fixture {name}", flush=True)
    gen, nkr, nocc = regenerate(g, tmp_path)
    r = RelCCSD(nocc, 2 * nkr - nocc)
    e = r.read_molfdir(str(tmp_path), [nocc])
    for c in CLASSES:
        assert np.array_equal(r.get(c), g[c]), c
    for k in ("FOO", "FVO", "FVV"):
        assert np.abs(r.get(k) - g[k]).max() < 1e-14, k
    assert np.abs(r.get_eps() - g["eps"]).max() < 1e-14
    assert abs(e["e0"] - float(g["E0"])) < 1e-12 and abs(e["e1"] - float(g["E1"])) < 1e-12
    assert abs(e["escf"] - float(g["ESCF"])) < 1e-12
    assert r.last_sort_ms() >= 0.0


@pytest.mark.parametrize("name", names("sortcont"))
def test_read_molfdir_follows_continuation_files(name, tmp_path):
    """The list spread over MDCINT, MDCINT1, MDCINT2 (continuation records)
    gives the single-file reference sort bit for bit, every class; the
    classes relccsd's first pass sorts also equal its own continuation run
    (its VOVV / VVVV passes re-read only the last file; see
    test_oracle_rdints.test_reference_reads_continuation_files)."""
    from molfdir_amd import mdcint
    g = load("sortcont", name)
    single = load("sort", name)
    nkr, nocc, seed = (int(x) for x in g["params"])
    print(f"sortcont fixture {name}", flush=True)
    mdcint.synthetic_files(str(tmp_path), nkr, nocc, seed, float(g["keep"]), parts=int(g["parts"]))
    r = RelCCSD(nocc, 2 * nkr - nocc)
    e = r.read_molfdir(str(tmp_path), [nocc])
    for c in CLASSES:
        assert np.array_equal(r.get(c), single[c]), c
    for c in ("OOOO", "VOOO", "VVOO", "VOVO"):
        assert np.array_equal(r.get(c), g[c]), c
    assert abs(e["e0"] - float(g["E0"])) < 1e-12 and abs(e["e1"] - float(g["E1"])) < 1e-12
    (tmp_path / "MDCINT2").unlink()
    with pytest.raises(RccError):
        r.read_molfdir(str(tmp_path), [nocc])


@pytest.mark.parametrize("name", names("sortbreit"))
def test_read_molfdir_adds_the_breit_list(name, tmp_path):
    """BREIT in MRCONEE: the MDBINT list is sorted with ITRSGN = -1 and the
    Coulomb classes are added, bit for bit as relccsd's RDINTS; a missing
    MDBINT is an error."""
    from molfdir_amd import mdcint
    g = load("sortbreit", name)
    nkr, nocc, seed = (int(x) for x in g["params"])
    print(f"sortbreit fixture {name}", flush=True)
    mdcint.synthetic_files(str(tmp_path), nkr, nocc, seed, float(g["keep"]), breit=True)
    r = RelCCSD(nocc, 2 * nkr - nocc)
    e = r.read_molfdir(str(tmp_path), [nocc])
    for c in CLASSES:
        assert np.array_equal(r.get(c), g[c]), c
    for k in ("FOO", "FVO", "FVV"):
        assert np.abs(r.get(k) - g[k]).max() < 1e-14, k
    assert abs(e["e0"] - float(g["E0"])) < 1e-12 and abs(e["e1"] - float(g["E1"])) < 1e-12
    (tmp_path / "MDBINT").unlink()
    with pytest.raises(RccError):
        r.read_molfdir(str(tmp_path), [nocc])


@pytest.mark.parametrize("name", names("sortopt"))
def test_read_molfdir_ccsort_options(name, tmp_path):
    """Namelist CCSORT's USEOE / NORECMP through rcc_read_molfdir_opts: the
    Fock blocks, orbital energies and E0 / E1 relccsd's GETFM leaves."""
    from molfdir_amd import mdcint
    g = load("sortopt", name)
    nkr, nocc, seed, useoe, norecmp = (int(x) for x in g["params"])
    print(f"sortopt fixture {name}", flush=True)
    mdcint.synthetic_files(str(tmp_path), nkr, nocc, seed, float(g["keep"]))
    r = RelCCSD(nocc, 2 * nkr - nocc)
    e = r.read_molfdir(str(tmp_path), [nocc], useoe=bool(useoe), norecmp=bool(norecmp))
    for c in CLASSES:
        assert np.array_equal(r.get(c), g[c]), c
    for k in ("FOO", "FVO", "FVV"):
        assert np.abs(r.get(k) - g[k]).max() < 1e-14, k
    assert np.abs(r.get_eps() - g["eps"]).max() < 1e-14
    assert abs(e["e0"] - float(g["E0"])) < 1e-12 and abs(e["e1"] - float(g["E1"])) < 1e-12


def test_device_list_entry_point_gives_the_same_classes(tmp_path):
    g = load("sort", "k9_5o")
    gen, nkr, nocc = regenerate(g, tmp_path)
    nsp = 2 * nkr
    r = RelCCSD(nocc, nsp - nocc)
    pos = np.concatenate([np.arange(nocc), -(np.arange(nsp - nocc) + 1)]).astype(np.int32)
    idx = torch.tensor(np.stack([gen["ikr"], gen["jkr"], gen["kkr"], gen["lkr"]], 1),
                       dtype=torch.int32, device="cuda")
    val = torch.tensor(gen["val"], dtype=torch.complex128, device="cuda")
    ms = r.sort_ints_device(pos, gen["kr"], idx, val)
    assert ms >= 0.0
    for c in CLASSES:
        assert np.array_equal(r.get(c), g[c]), c


def test_inconsistent_inputs_are_rejected(tmp_path):
    g = load("sort", "k6_4o")
    gen, nkr, nocc = regenerate(g, tmp_path)
    r = RelCCSD(nocc + 1, 2 * nkr - nocc - 1)
    with pytest.raises(RccError):
        r.read_molfdir(str(tmp_path), [nocc])          # handle of another occupation
    r2 = RelCCSD(nocc, 2 * nkr - nocc)
    with pytest.raises(RccError):
        r2.read_molfdir(str(tmp_path), [nocc, 0])      # two NELEC values, one irrep
    with pytest.raises(RccError):
        r2.read_molfdir(str(tmp_path / "missing"), [nocc])


@pytest.mark.parametrize("record", [2, 3])
def test_truncated_mrconee_records_are_rejected(record, tmp_path):
    """An MRCONEE whose record 2 (NSYMRP) or 3 (NREP) is shorter than the
    4-byte count RDSYMI reads first (ccdriv.F:2978-2988) is refused with an
    error, not read past its end."""
    import struct
    g = load("sort", "k6_4o")
    gen, nkr, nocc = regenerate(g, tmp_path)
    raw = (tmp_path / "MRCONEE").read_bytes()
    recs, p = [], 0
    while p < len(raw):
        n = struct.unpack_from("<i", raw, p)[0]
        recs.append(raw[p + 4:p + 4 + n])
        p += n + 8
    recs[record - 1] = b"\x07\x00"
    (tmp_path / "MRCONEE").write_bytes(b"".join(
        struct.pack("<i", len(x)) + x + struct.pack("<i", len(x)) for x in recs))
    r = RelCCSD(nocc, 2 * nkr - nocc)
    with pytest.raises(RccError, match=f"record {record}"):
        r.read_molfdir(str(tmp_path), [nocc])


@pytest.mark.parametrize("kind,name,opts", [("sort", "k7_6o", 0),
                                             ("sortopt", "k7_6o_norecmp", 2),
                                             ("sortopt", "k7_6o_useoe", 1)])
def test_fortran_driver_sorts_molfdir_files(kind, name, opts, tmp_path):
    # JOB 4 of the Fortran driver: rcc_read_molfdir_opts through ISO_C_BINDING,
    # header word 8 carrying namelist CCSORT's options
    import os
    import subprocess
    from molfdir_amd import mdcint
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "molfdir_amd", "fortran", "ccdrv_hip.x")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built")
    g = load(kind, name)
    nkr, nocc, seed = (int(x) for x in g["params"][:3])
    mdcint.synthetic_files(str(tmp_path), nkr, nocc, seed, float(g["keep"]))
    nv = 2 * nkr - nocc
    hdr = np.array([nocc, nv, 4, 1, 1, 12, 0, opts], dtype=np.int32)
    with open(tmp_path / "ccin.bin", "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.zeros(nocc + nv).tobytes())
    p = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    raw = np.fromfile(tmp_path / "ccout.bin", dtype=np.float64)
    assert abs(raw[0] - float(g["E0"])) < 1e-12 and abs(raw[1] - float(g["E1"])) < 1e-12
    pos = 3
    for k in ("FOO", "FVO", "FVV") + CLASSES:
        n = g[k].size
        x = raw[pos:pos + 2 * n].view(np.complex128)
        pos += 2 * n
        if k in CLASSES:
            assert np.array_equal(x, g[k]), k
        else:
            assert np.abs(x - g[k]).max() < 1e-14, k
    assert np.abs(raw[pos:] - g["eps"]).max() < 1e-14


@pytest.mark.parametrize("name", ["cfg1_6o40v", "s4_5o12v"])
def test_handle_restart_file_round_trip(name, tmp_path):
    # the handle reads relccsd's MCCRES into T and writes it back unchanged
    import os
    from _golden import GOLDEN
    m = load("mccres", name)
    kind = "ccsd" if name.startswith("cfg") else "symccsd"
    g = load(kind, name)
    no, nv = int(g["params"][0]), int(g["params"][1])
    sym = (g["sym_no"].tolist(), g["sym_nv"].tolist()) if "sym_no" in g else None
    r = RelCCSD(no, nv, sym=sym)
    st, en = r.read_mccres(os.path.join(GOLDEN, f"mccres_{name}.bin"))
    t1, t2 = r.get_amplitudes()
    assert np.array_equal(t1, m["T1"]) and np.array_equal(t2, m["T2"])
    assert np.array_equal(st, m["IMSTAT"]) and np.array_equal(en, m["energies"])
    out = tmp_path / "MCCRES"
    r.write_mccres(out, imstat=st, energies=en)
    from molfdir_amd.relccsd import mccres_read
    a1, a2, st2, en2 = mccres_read(out, m["T1"].size, m["T2"].size)
    assert np.array_equal(a1, m["T1"]) and np.array_equal(a2, m["T2"])


def test_read_molfdir_symmetric_breit_matches_reference_sort(tmp_path):
    # a symmetry-blocked handle with the Breit list (BREIT flag, MDBINT): the
    # Breit pass also sorts into the C1-sized VVVV scratch before the irrep
    # blocks are gathered
    from test_oracle_rdints import regenerate_sym
    for name in names("symsortbreit"):
        g = load("symsortbreit", name)
        print(f"symsortbreit fixture {name}", flush=True)
        gen = regenerate_sym(g, tmp_path, breit=True)
        no_l, nv_l = g["sym_no"].tolist(), g["sym_nv"].tolist()
        r = RelCCSD(sum(no_l), sum(nv_l), sym=(no_l, nv_l))
        e = r.read_molfdir(str(tmp_path), gen["nelec"].tolist())
        for c in CLASSES:
            assert np.array_equal(r.get(c), g[c]), c
        for k in ("FOO", "FVO", "FVV"):
            assert np.abs(r.get(k) - g[k]).max() < 1e-14, k
        assert abs(e["e0"] - float(g["E0"])) < 1e-12 and abs(e["e1"] - float(g["E1"])) < 1e-12


@pytest.mark.parametrize("name", names("symsort"))
def test_read_molfdir_symmetric_matches_reference_sort(name, tmp_path):
    # a symmetry-blocked handle: the classes (VOVV / VVVV in their irrep
    # blocks) and the Fock blocks equal relccsd's symmetric sort
    from test_oracle_rdints import regenerate_sym
    g = load("symsort", name)
    print(f"symsort fixture {name}", flush=True)
    gen = regenerate_sym(g, tmp_path)
    no_l, nv_l = g["sym_no"].tolist(), g["sym_nv"].tolist()
    r = RelCCSD(sum(no_l), sum(nv_l), sym=(no_l, nv_l))
    e = r.read_molfdir(str(tmp_path), gen["nelec"].tolist())
    for c in CLASSES:
        assert np.array_equal(r.get(c), g[c]), c
    for k in ("FOO", "FVO", "FVV"):
        assert np.abs(r.get(k) - g[k]).max() < 1e-14, k
    assert np.abs(r.get_eps() - g["eps"]).max() < 1e-14
    assert abs(e["e0"] - float(g["E0"])) < 1e-12 and abs(e["e1"] - float(g["E1"])) < 1e-12


@pytest.mark.parametrize("kind", ["sort", "sortbreit", "symsort"])
def test_anchor_keyed_sort_matches_reference_sort(kind, tmp_path, monkeypatch):
    """The sort's second key form (RCC_SORT_KEYS=anchor: each orbit keyed by
    its smallest stored expansion, in the class memory, with a flag pass),
    which lists whose canonical key array would pass 4 GiB use: the same
    classes bit for bit on every fixture of the kind."""
    from test_oracle_rdints import regenerate_sym
    monkeypatch.setenv("RCC_SORT_KEYS", "anchor")
    for name in names(kind):
        g = load(kind, name)
        print(f"anchor-keyed sort, {kind} fixture {name}", flush=True)
        d = tmp_path / name
        d.mkdir()
        if kind == "symsort":
            gen = regenerate_sym(g, d)
            no_l, nv_l = g["sym_no"].tolist(), g["sym_nv"].tolist()
            r = RelCCSD(sum(no_l), sum(nv_l), sym=(no_l, nv_l))
            r.read_molfdir(str(d), gen["nelec"].tolist())
        else:
            from molfdir_amd import mdcint
            nkr, nocc, seed = (int(x) for x in g["params"])
            mdcint.synthetic_files(str(d), nkr, nocc, seed, float(g["keep"]),
                                   breit=(kind == "sortbreit"))
            r = RelCCSD(nocc, 2 * nkr - nocc)
            r.read_molfdir(str(d), [nocc])
        for c in CLASSES:
            assert np.array_equal(r.get(c), g[c]), (name, c)
        r.close()
