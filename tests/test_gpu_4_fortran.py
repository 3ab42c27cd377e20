"""The Fortran driver (ISO_C_BINDING over the C ABI, molfdir_amd/fortran)
reproduces the relccsd reference trajectory on the GPU."""
import os
import subprocess

import numpy as np
import pytest

from oracle import refrun

from _golden import load, names, system_for

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "molfdir_amd", "fortran", "ccdrv_hip.x")


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
@pytest.mark.parametrize("kind,name", [("ccsd", "cfg1_6o40v"), ("ccsd", "conv_5o16v"),
                                       ("symccsd", "s4_5o12v"), ("symccsd", "s8_8o24v"),
                                       ("frzccsd", "c1_6o40v"), ("frzccsd", "s4_7o16v")])
def test_fortran_driver_matches_reference(kind, name, tmp_path):
    g = load(kind, name)
    no, nv, seed, maxit, maxdim, ntol = (int(x) for x in g["params"])
    s = system_for(g)
    refrun._write_input(str(tmp_path / "ccin.bin"), s, 1, maxit, maxdim, ntol)
    p = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    rows = refrun.parse_iterations(p.stdout)
    ref = g["iterations"]
    assert len(rows) == len(ref)
    for a, b in zip(rows[1:], ref[1:]):
        assert int(a[0]) == int(b[0])
        assert abs(a[1] - b[1]) < 1e-10
        assert abs(a[2] - b[2]) < 1e-9 * max(1.0, b[2])
        # T1-DIAGN: both tables print it F10.5; one unit of the last digit
        # covers a value that sits on a rounding boundary
        assert abs(a[3] - b[3]) <= 1.0000001e-5, (a, b)
    raw = np.fromfile(tmp_path / "ccout.bin", dtype=np.float64)
    z = raw[3:].view(np.complex128)
    nt1 = g["T1"].size
    assert z.size == nt1 + g["T2"].size
    assert abs(raw[0] - float(g["emp2"])) < 1e-10
    assert np.abs(z[:nt1] - g["T1"]).max() < 1e-9
    assert np.abs(z[nt1:] - g["T2"]).max() < 1e-9
    # the restart file the Fortran driver leaves, as relccsd's CCENER leaves MCCRES
    from molfdir_amd.relccsd import mccres_read
    t1, t2, st, en = mccres_read(tmp_path / "MCCRES", nt1, g["T2"].size)
    assert np.array_equal(t1, z[:nt1]) and np.array_equal(t2, z[nt1:])
    if name in names("mccres"):
        m = load("mccres", name)
        assert np.array_equal(st, m["IMSTAT"])
        assert np.abs(en - m["energies"]).max() < 1e-10


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_driver_lambda_density_gives_reference_gradient(tmp_path):
    g = load("lambda", "lam_4o12v")
    no, nv = int(g["params"][0]), int(g["params"][1])
    s = system_for(g)
    path = tmp_path / "ccin.bin"
    refrun._write_input(str(path), s, 1, 80, 8, 12)
    raw = bytearray(path.read_bytes())
    raw[28:32] = np.array([1], dtype=np.int32).tobytes()     # header word 8: LAMBDA
    path.write_bytes(bytes(raw))
    p = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    out = np.fromfile(tmp_path / "ccout.bin", dtype=np.float64)
    nt1, nt2 = no * nv, (no * (no - 1) // 2) * (nv * (nv - 1) // 2)
    pos = 3 + 2 * (nt1 + nt2)
    lnit, lrms = out[pos], out[pos + 1]
    n = no + nv
    D = out[pos + 2:].view(np.complex128).reshape(n, n, order="F")
    assert lrms < 1e-12 and lnit > 0
    grad = float(np.real(np.trace(D @ g["V"])))
    fd = float(g["fd4"])
    assert abs(grad - fd) <= 2e-6 * abs(fd) + 2e-11, (grad, fd)


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_driver_triples_match_reference(tmp_path):
    # JOB 2: CCSD, then T3CORR on the converged amplitudes (DOCCSDT)
    g = load("ccsd", "cfg1_6o40v")
    t = load("t3", "conv_6o40v")
    no, nv, seed, maxit, maxdim, ntol = (int(x) for x in g["params"])
    s = system_for(g)
    path = tmp_path / "ccin.bin"
    refrun._write_input(str(path), s, 1, maxit, maxdim, ntol)
    raw = bytearray(path.read_bytes())
    raw[8:12] = np.array([2], dtype=np.int32).tobytes()      # header word 3: JOB
    path.write_bytes(bytes(raw))
    p = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "Total CCSD(T) energy" in p.stdout
    out = np.fromfile(tmp_path / "ccout.bin", dtype=np.float64)
    nt1, nt2 = g["T1"].size, g["T2"].size
    et = out[3 + 2 * (nt1 + nt2):]
    assert et.size == 3
    for x, key in zip(et, ("ET1", "ET2", "ET3")):
        assert abs(x - float(t[key])) < 1e-10, (key, x, float(t[key]))


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_driver_mp2_relaxed_property_matches_reference(tmp_path):
    # JOB 5: CCFOPR's MP2 block (MP2 lambda, CCDENSC, CCDENSZ + RPAA, EXPVAL)
    # on the operators named at the end of ccin.bin, read from MDPROP
    from molfdir_amd import mdcint, synth
    g = load("mp2g", "g_6o20v")
    no, nv, _, pseed, maxdim, ntol = (int(x) for x in g["params"])
    s = system_for(g)
    P = synth.hermitian_perturbation(no + nv, pseed, 0.1)
    eps = s["eps"].numpy()
    mdcint.write_mrconee(str(tmp_path / "MRCONEE"), eps, np.ones(no + nv, np.int32), [no],
                         np.diag(eps).astype(np.complex128))
    mdcint.write_mdprop(str(tmp_path / "MDPROP"), {"OTHER": 2.0 * P, "PROPTEST": P})
    path = tmp_path / "ccin.bin"
    refrun._write_input(str(path), s, 1, 100, maxdim, ntol)
    raw = bytearray(path.read_bytes())
    raw[8:12] = np.array([5], dtype=np.int32).tobytes()      # header word 3: JOB
    raw += np.array([2], dtype=np.int32).tobytes() + b"PROPTESTOTHER   "
    path.write_bytes(bytes(raw))
    p = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "MP2 Gradient results" in p.stdout and "<PROPTEST>" in p.stdout
    out = np.fromfile(tmp_path / "ccout.bin", dtype=np.float64)
    assert out.size == 4 * 2 + 2
    scf, mp2 = complex(out[0], out[1]), complex(out[2], out[3])
    assert abs(scf - np.trace(P[:no, :no])) < 1e-14
    assert abs(mp2 - complex(g["expval"])) < 1e-11, (mp2, g["expval"])
    # the second operator is 2 P: both values double (EXPVAL is linear)
    assert abs(complex(out[4], out[5]) - 2 * scf) < 1e-13
    assert abs(complex(out[6], out[7]) - 2 * mp2) < 1e-11
    assert out[-1] < 10.0 ** -ntol
