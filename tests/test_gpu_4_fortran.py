"""The Fortran side on the GPU.

* The Fortran driver (ISO_C_BINDING over the C ABI, molfdir_amd/fortran/
  ccdrv_hip.f90) reproduces the relccsd reference trajectory.
* relccsd's own driver as the caller: the relccsd objects built from the
  reference sources with molfdir_amd/fortran/cceqn_hip.F linked in place of
  cceqns.F's MP2EQN and CCEQN (oracle/_ref/cchrns_gpu.x, whose CCENER runs
  MP2 + CCSD; oracle/_ref/relccsd_gpu.x, the whole PROGRAM RELCCSD with its
  own integral sort of MOLFDIR files) print the CPU reference's MP2 energy,
  iteration table and CCSD energy and leave its amplitudes and MCCRES
  records."""
import os
import subprocess

import numpy as np
import pytest

from oracle import refrun

from _golden import load, names, system_for

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "molfdir_amd", "fortran", "ccdrv_hip.x")
CASES = [("ccsd", "cfg1_6o40v"), ("ccsd", "conv_5o16v"), ("symccsd", "s4_5o12v"),
         ("symccsd", "s8_8o24v"), ("frzccsd", "c1_6o40v"), ("frzccsd", "s4_7o16v")]


def check_rows(rows, ref):
    assert len(rows) == len(ref)
    for a, b in zip(rows[1:], ref[1:]):
        assert int(a[0]) == int(b[0])
        assert abs(a[1] - b[1]) < 1e-10, (a, b)
        assert abs(a[2] - b[2]) < 1e-9 * max(1.0, b[2]), (a, b)
        # T1-DIAGN: both tables print it F10.5; one unit of the last digit
        # covers a value that sits on a rounding boundary
        assert abs(a[3] - b[3]) <= 1.0000001e-5, (a, b)


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
@pytest.mark.parametrize("kind,name", CASES)
def test_fortran_driver_matches_reference(kind, name, tmp_path):
    g = load(kind, name)
    no, nv, seed, maxit, maxdim, ntol = (int(x) for x in g["params"])
    s = system_for(g)
    refrun._write_input(str(tmp_path / "ccin.bin"), s, 1, maxit, maxdim, ntol)
    p = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    check_rows(refrun.parse_iterations(p.stdout), g["iterations"])
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


@pytest.mark.skipif(not os.path.exists(refrun.EXE_GPU), reason="GPU-linked relccsd not built")
@pytest.mark.parametrize("kind,name", CASES)
def test_relccsd_ccener_drives_the_gpu_solver(kind, name, tmp_path):
    """relccsd's CCENER (ccdriv.F:1715-1792) calling the GPU MP2EQN / CCEQN."""
    g = load(kind, name)
    no, nv, seed, maxit, maxdim, ntol = (int(x) for x in g["params"])
    print(f"CCENER -> GPU solver: {kind} {name}", flush=True)
    r = refrun.run(system_for(g), "ccsd", maxit=maxit, maxdim=maxdim, ntol=ntol,
                   exe=refrun.EXE_GPU, workdir=str(tmp_path), timeout=300)
    assert "Timing of routine CCEQN (GPU solver" in r["log"]
    check_rows(r["iterations"], g["iterations"])
    assert abs(r["emp2"] - float(g["emp2"])) < 1e-10
    assert abs(r["eccsd"] - float(g["eccsd"])) < 1e-10
    assert r["nit"] == int(g["nit"])
    # what CCENER's PUTAMPT wrote to MCCRES after CCEQN, read back by GETAMPT
    assert np.abs(r["T1"] - g["T1"]).max() < 1e-9
    assert np.abs(r["T2"] - g["T2"]).max() < 1e-9


@pytest.mark.skipif(not os.path.exists(refrun.PROGRAM_GPU), reason="GPU-linked relccsd not built")
@pytest.mark.parametrize("name", names("prog"))
def test_relccsd_program_with_the_gpu_solver(name, tmp_path):
    """PROGRAM RELCCSD: its own integral sort of MOLFDIR files (CCSETI ->
    RDINTS), then CCENER's MP2 and CCSD on the GPU solver."""
    from molfdir_amd import mdcint
    from molfdir_amd.relccsd import mccres_read
    g = load("prog", name)
    seed, maxit, maxdim, ntol = (int(x) for x in g["params"])
    nkr, nelec = g["nkr"].tolist(), g["nelec"].tolist()
    nfroz = g["nfroz"].tolist() or None
    print(f"RELCCSD -> GPU solver: {name}", flush=True)
    if len(nkr) == 1:
        mdcint.synthetic_files(str(tmp_path), nkr[0], nelec[0], seed, float(g["keep"]),
                               scale=float(g["scale"]))
    else:
        mdcint.synthetic_files_sym(str(tmp_path), nkr, nelec, seed, float(g["keep"]),
                                   scale=float(g["scale"]))
    r = refrun.run_program(str(tmp_path), nelec, maxit, maxdim, ntol, nfroz=nfroz,
                           exe=refrun.PROGRAM_GPU, timeout=300, keep_dir=str(tmp_path))
    assert "Timing of routine CCEQN (GPU solver" in r["log"]
    # the sort is relccsd's own in both runs
    assert abs(r["escf"] - float(g["escf"])) < 1e-12
    assert abs(r["emp2"] - float(g["emp2"])) < 1e-10
    assert abs(r["eccsd"] - float(g["eccsd"])) < 1e-10
    check_rows(r["iterations"], g["iterations"])
    t1, t2, st, en = mccres_read(tmp_path / "MCCRES", g["T1"].size, g["T2"].size)
    assert np.abs(t1 - g["T1"]).max() < 1e-9
    assert np.abs(t2 - g["T2"]).max() < 1e-9
    assert np.array_equal(st, g["IMSTAT"])
    assert np.abs(en - g["energies"]).max() < 1e-10
