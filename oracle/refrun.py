"""Runs the compiled relccsd reference (oracle/_ref/cchrns.x) on a system.

Test infrastructure only: used by tests/ and tests/golden/ scripts to obtain
reference outputs.  The executable is built by ``make -f Makefile.ref`` from
the unmodified relccsd sources plus harness/cchrns.F.

Mode 1 ("ccsd") runs relccsd's MP2 + CCSD driver CCENER (ccdriv.F:1574)
with the given MAXIT / MAXDIM / NTOL and returns the final amplitudes and
the per-iteration table CCEQN prints (cceqns.F:270-272, 550-551).
Mode 2 ("terms") evaluates every CCEQN term once on given T1/T2 and returns
each intermediate (HO, HV, HOV, S1, GO, GV, AZ, H, S2 before and after BINTM,
S1/S2 after DENOM and ZCORE, CCENG's energy and T1DIAG of the new T1).
A system dict may carry "nfroz": frozen occupied spinors per irrep (namelist
RELCCSD NFROZ, SYMTAB ccdriv.F:838-849), the lowest of each irrep in file
order.
Mode 4 (run_sort) sorts a MOLFDIR MRCONEE + MDCINT pair with relccsd's own
CCSETI -> RDINTS (ccints.F:33-569): the six classes, the Fock blocks after
GETFM / MAKEFM, the orbital energies and ESCF.
Mode 3 ("t3") evaluates the non-iterative triples corrections of T3CORR
(cceqns.F:1336-1863) on given T1/T2: ET1 (fourth-order +T), ET2 ((T)) and
ET3 (-T).
"""
from __future__ import annotations

import os
import re
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "_ref", "cchrns.x")
# relccsd's own PROGRAM RELCCSD (ccmain.F), CPU
PROGRAM = os.path.join(HERE, "_ref", "relccsd.x")
# the same two programs with the GPU solver linked in place of cceqns.F's
# MP2EQN and CCEQN (molfdir_amd/fortran/cceqn_hip.F): relccsd's CCENER is
# then the caller of the GPU solver
EXE_GPU = os.path.join(HERE, "_ref", "cchrns_gpu.x")
PROGRAM_GPU = os.path.join(HERE, "_ref", "relccsd_gpu.x")

CLASS_ORDER = ("FOO", "FVO", "FVV", "OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV")


def available() -> bool:
    return os.path.exists(EXE)


def _c(x):
    x = np.asarray(x.detach().cpu().numpy() if hasattr(x, "detach") else x)
    return np.ascontiguousarray(x.astype(np.complex128).reshape(-1))


def layout_of(sysd):
    """The symmetry layout of a system: sysd['sym'] = (no per irrep, nv per
    irrep) for an abelian double group of xor_table type, else C1."""
    from oracle.symtab import SymLayout
    if "sym" in sysd and sysd["sym"] is not None:
        no_l, nv_l = sysd["sym"]
        return SymLayout(no_l, nv_l)
    return SymLayout([sysd["no"]], [sysd["nv"]])


def _write_input(path, sysd, mode, maxit, maxdim, ntol, t1=None, t2=None):
    """ccin.bin for harness/cchrns.F.  Classes (and t1/t2) are given in C1
    layout and written in the reference's symmetry layout."""
    from oracle.symtab import multb_full
    lay = layout_of(sysd)
    no, nv = sysd["no"], sysd["nv"]
    nrep = lay.nrep
    nfroz = sysd.get("nfroz")
    hdr = np.array([no, nv, mode, maxit, maxdim, ntol, nrep if nrep > 1 else 0,
                    2 if nfroz is not None and any(nfroz) else 0], dtype=np.int32)
    eps = np.asarray(sysd["eps"].cpu().numpy() if hasattr(sysd["eps"], "cpu") else sysd["eps"],
                     dtype=np.float64)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(eps.tobytes())
        if nrep > 1:
            irr = np.concatenate([lay.occ_irreps(), lay.vir_irreps()]) + 1
            f.write(irr.astype(np.int32).tobytes())
            f.write(lay.no.astype(np.int32).tobytes())
            f.write(multb_full(nrep).T.astype(np.int32).tobytes())   # column-major
        if hdr[7] & 2:
            f.write(np.asarray(nfroz, dtype=np.int32).reshape(nrep).tobytes())
        for name in CLASS_ORDER:
            f.write(_c(lay.to_sym(name, _c(sysd[name]))).tobytes())
        if mode in (2, 3):
            f.write(_c(lay.to_sym("T1", _c(t1))).tobytes())
            f.write(_c(lay.to_sym("T2", _c(t2))).tobytes())


_ITER_RE = re.compile(r"^\s+(\d+)\s+(-?[\d.]+(?:[EeDd][-+]?\d+)?)\s+(-?[\d.]+(?:[EeDd][-+]?\d+)?)\s+(-?[\d.]+)\s*$")


def parse_iterations(log: str):
    """The CCEQN table rows 'NIT ENERGY RMS T1-DIAGN' (format 6001)."""
    rows = []
    in_table = False
    for line in log.splitlines():
        if "NIT" in line and "ENERGY" in line and "RMS" in line:
            in_table = True
            rows = []
            continue
        if in_table:
            m = _ITER_RE.match(line)
            if m:
                rows.append(tuple(float(g.replace("D", "E")) for g in m.groups()))
            elif "CCSD results" in line or "Requested convergence" in line:
                in_table = False
    return rows


def run(sysd, mode="ccsd", maxit=20, maxdim=6, ntol=12, t1=None, t2=None,
        workdir=None, timeout=3600, keep_mccres=None, exe=None):
    """Outputs are in the reference's (symmetry) layout; for a C1 system
    that is the C1 layout.  exe: the harness binary (default the CPU one)."""
    lay = layout_of(sysd)
    nt1, nt2 = lay.size("T1"), lay.size("T2")
    m = {"ccsd": 1, "terms": 2, "t3": 3, "restart": 5}[mode]
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        _write_input(os.path.join(d, "ccin.bin"), sysd, m, maxit, maxdim, ntol, t1, t2)
        p = subprocess.run([exe or EXE], cwd=d, capture_output=True, text=True,
                           timeout=timeout)
        log = p.stdout + p.stderr
        if p.returncode != 0 or not os.path.exists(os.path.join(d, "ccout.bin")):
            raise RuntimeError(f"reference harness failed (rc={p.returncode}):\n{log[-4000:]}")
        raw = np.fromfile(os.path.join(d, "ccout.bin"), dtype=np.float64)
        if keep_mccres:
            import shutil
            shutil.copy(os.path.join(d, "MCCRES"), keep_mccres)
    out = {"log": log}
    if m == 1:
        out["emp2"], out["eccsd"], out["nit"] = raw[0], raw[1], int(raw[2])
        z = raw[3:].view(np.complex128)
        out["T1"], out["T2"] = z[:nt1].copy(), z[nt1:nt1 + nt2].copy()
        out["iterations"] = parse_iterations(log)
        mt = re.search(r"Total time in CCEQN :\s+([\d.]+)", log)
        out["cceqn_cpu_s"] = float(mt.group(1)) if mt else None
        return out
    if m == 3:
        out["ET1"], out["ET2"], out["ET3"] = raw[0], raw[1], raw[2]
        mt = re.search(r"Total time in T3CORR :\s+([\d.]+)", log)
        out["t3corr_cpu_s"] = float(mt.group(1)) if mt else None
        return out
    nfoo, nfvo, nfvv = lay.size("FOO"), lay.size("FVO"), lay.size("FVV")
    nv1, nv4 = lay.size("OOOO"), lay.size("VOVO")
    sizes = [("HO", nfoo), ("HV", nfvv), ("HOV", nfvo), ("S1", nt1), ("GO", nfoo),
             ("GV", nfvv), ("AZ", nv1), ("H", nv4), ("S2_T2EQNS", nt2),
             ("S2_BINTM", nt2), ("S1_DENOM", nt1), ("S2_DENOM", nt2)]
    z = raw[:-2].view(np.complex128)
    pos = 0
    for name, n in sizes:
        out[name] = z[pos:pos + n].copy()
        pos += n
    assert pos == z.size, (pos, z.size)
    out["ECC"], out["T1DIAG"] = raw[-2], raw[-1]
    return out


def run_sort(files_dir, no, nv, workdir=None, timeout=3600, sym=None, useoe=False,
             norecmp=False):
    """Integral sort of files_dir/MRCONEE + files_dir/MDCINT.  sym (for
    NREP > 1): (file-order irrep of every spinor (1-based), occupied count
    per irrep, nv per irrep); outputs are then in the symmetry-blocked layout.
    useoe / norecmp: namelist CCSORT's USEOE and NORECMP (GETFM ccints.F:
    1080, 1150-1198, 1260, 1287)."""
    import shutil
    from oracle.symtab import SymLayout, multb_full
    if sym is None:
        lay = SymLayout([no], [nv])
    else:
        lay = SymLayout(list(sym[1]), list(sym[2]))
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        # MRCONEE, MDCINT, any continuation files MDCINT1, MDCINT2, .. and the
        # Breit list MDBINT
        for fn in sorted(os.listdir(files_dir)):
            if fn == "MRCONEE" or fn.startswith("MDCINT") or fn.startswith("MDBINT"):
                shutil.copy(os.path.join(files_dir, fn), os.path.join(d, fn))
        nrep = lay.nrep
        opts = (1 if useoe else 0) + (2 if norecmp else 0)
        hdr = np.array([no, nv, 4, 1, 1, 12, nrep if nrep > 1 else 0, opts], dtype=np.int32)
        with open(os.path.join(d, "ccin.bin"), "wb") as f:
            f.write(hdr.tobytes())
            f.write(np.zeros(no + nv).tobytes())
            if nrep > 1:
                f.write(np.asarray(sym[0], dtype=np.int32).tobytes())
                f.write(np.asarray(sym[1], dtype=np.int32).tobytes())
                f.write(multb_full(nrep).T.astype(np.int32).tobytes())
        p = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=timeout)
        log = p.stdout + p.stderr
        if p.returncode != 0 or not os.path.exists(os.path.join(d, "ccout.bin")):
            raise RuntimeError(f"reference harness failed (rc={p.returncode}):\n{log[-4000:]}")
        raw = np.fromfile(os.path.join(d, "ccout.bin"), dtype=np.float64)
    out = {"log": log}
    pos = 0
    for name in ("FOO", "FVO", "FVV"):
        n = lay.size(name)
        out[name] = raw[pos:pos + 2 * n].view(np.complex128).copy()
        pos += 2 * n
    out["eps"] = raw[pos:pos + no + nv].copy()
    pos += no + nv
    for name in ("OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV"):
        n = lay.size(name)
        out[name] = raw[pos:pos + 2 * n].view(np.complex128).copy()
        pos += 2 * n
    out["ESCF"] = raw[pos]
    assert pos + 1 == raw.size, (pos, raw.size)
    for key, pat in (("E0", r"Zero order electronic energy :\s+(-?[\d.]+)"),
                     ("E1", r"First order electronic energy :\s+(-?[\d.]+)")):
        m = re.search(pat, log)
        out[key] = float(m.group(1)) if m else None
    m = re.search(r"Total time in .*?sort.*?:\s+([\d.]+)", log, re.I)
    return out


def run_program(files_dir, nelec, maxit=20, maxdim=6, ntol=10, nfroz=None, exe=None,
                timeout=3600, keep_dir=None):
    """relccsd itself (PROGRAM RELCCSD, ccmain.F) on files_dir/MRCONEE +
    MDCINT: the integral sort CCSETI -> RDINTS, then CCENER's MP2 and CCSD,
    from a MOLFDIR.INP with namelists RELCCSD (NELEC and NFROZ per irrep,
    CARITH, DOSORT, DOENER, IPRNT = 1, TIMING) and CCENER (MAXIT, MAXDIM,
    NTOL; ccdriv.F:155-225).  Returns the log, the printed SCF / MP2 / CCSD
    energies, the CCEQN table and the MCCRES restart file's bytes.  exe:
    the program binary (default the CPU one)."""
    import shutil
    line = ",".join(str(int(x)) for x in nelec)
    frz = "" if nfroz is None else " NFROZ=" + ",".join(str(int(x)) for x in nfroz) + ","
    with tempfile.TemporaryDirectory() as d:
        for fn in sorted(os.listdir(files_dir)):
            if fn == "MRCONEE" or fn.startswith("MDCINT") or fn.startswith("MDBINT"):
                shutil.copy(os.path.join(files_dir, fn), os.path.join(d, fn))
        with open(os.path.join(d, "MOLFDIR.INP"), "w") as f:
            f.write(f" &RELCCSD NELEC={line},{frz} CARITH=T, DOSORT=T, DOENER=T, IPRNT=1,"
                    f" TIMING=T, MWORD=15 /\n")
            f.write(f" &CCENER DOMP2=T, DOCCSD=T, MAXIT={maxit}, MAXDIM={maxdim},"
                    f" NTOL={ntol} /\n")
        p = subprocess.run([exe or PROGRAM], cwd=d, capture_output=True, text=True,
                           timeout=timeout)
        log = p.stdout + p.stderr
        if p.returncode != 0 or "Normal end of RELCCSD" not in log:
            raise RuntimeError(f"relccsd failed (rc={p.returncode}):\n{log[-4000:]}")
        mccres = open(os.path.join(d, "MCCRES"), "rb").read()
        if keep_dir:
            shutil.copy(os.path.join(d, "MCCRES"), os.path.join(keep_dir, "MCCRES"))
    out = {"log": log, "iterations": parse_iterations(log), "mccres": mccres}
    for key, pat in (("escf", r"^ SCF energy :\s+(-?[\d.]+)"),
                     ("emp2", r"^ MP2 correlation energy :\s+(-?[\d.]+)"),
                     ("eccsd", r"^ CCSD correlation energy :\s+(-?[\d.]+)")):
        m = re.findall(pat, log, re.M)
        out[key] = float(m[-1]) if m else None
    return out


def read_restart(sysd, mccres_path, workdir=None, timeout=600):
    """Mode 5: relccsd's GETAMPT (and the CCDUMP records) on a given MCCRES
    file; returns T1, T2 (reference layout), IMSTAT and the 7 energies."""
    import shutil
    lay = layout_of(sysd)
    nt1, nt2 = lay.size("T1"), lay.size("T2")
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        _write_input(os.path.join(d, "ccin.bin"), sysd, 5, 1, 1, 12)
        shutil.copy(mccres_path, os.path.join(d, "MCCRES.in"))
        p = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=timeout)
        log = p.stdout + p.stderr
        if p.returncode != 0 or not os.path.exists(os.path.join(d, "ccout.bin")):
            raise RuntimeError(f"reference harness failed (rc={p.returncode}):\n{log[-4000:]}")
        raw = np.fromfile(os.path.join(d, "ccout.bin"), dtype=np.float64)
    z = raw[:2 * (nt1 + nt2)].view(np.complex128)
    return {"T1": z[:nt1].copy(), "T2": z[nt1:].copy(),
            "IMSTAT": raw[2 * (nt1 + nt2):2 * (nt1 + nt2) + 30].astype(np.int64),
            "energies": raw[2 * (nt1 + nt2) + 30:].copy()}


def read_property(sysd, mdprop_path, workdir=None, timeout=600):
    """Mode 6: relccsd's RDPROP of operator 'PROPTEST' (blocks AOO, AVV, AVO,
    AOV in the reference layout) and EXPVAL on zero densities (the
    reference-determinant expectation value)."""
    import shutil
    lay = layout_of(sysd)
    nfoo, nfvv, nfvo = lay.size("FOO"), lay.size("FVV"), lay.size("FVO")
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        _write_input(os.path.join(d, "ccin.bin"), sysd, 6, 1, 1, 12)
        shutil.copy(mdprop_path, os.path.join(d, "MDPROP"))
        p = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=timeout)
        log = p.stdout + p.stderr
        if p.returncode != 0 or not os.path.exists(os.path.join(d, "ccout.bin")):
            raise RuntimeError(f"reference harness failed (rc={p.returncode}):\n{log[-4000:]}")
        raw = np.fromfile(os.path.join(d, "ccout.bin"), dtype=np.float64)
    z = raw[:-2].view(np.complex128)
    out, pos = {}, 0
    for name, n in (("AOO", nfoo), ("AVV", nfvv), ("AVO", nfvo), ("AOV", nfvo)):
        out[name] = z[pos:pos + n].copy()
        pos += n
    out["expval"] = complex(raw[-2], raw[-1])
    return out


def mp2_property(sysd, mdprop_path, maxit=60, maxdim=8, ntol=13, workdir=None, timeout=600):
    """Mode 7: the MP2 relaxed property block of CCFOPR (ccdriv.F:2058-2090)
    on operator 'PROPTEST' of an MDPROP file (reference layouts; the VO
    vectors over all NDIMX pairs, totally symmetric block first).  Returns the MP2
    amplitudes, DOO / DVV after CCDENSC(1) (packed, reference layout), the
    X intermediate as CCDENSZ leaves it (-X), the Z-vector DVO / DOV that
    RPAA converged, its preconditioner DG, EXPVAL's value and the CPU seconds
    of the chain (cpu_s)."""
    import shutil
    lay = layout_of(sysd)
    nt1, nt2, nfoo, nfvv = lay.size("T1"), lay.size("T2"), lay.size("FOO"), lay.size("FVV")
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        _write_input(os.path.join(d, "ccin.bin"), sysd, 7, maxit, maxdim, ntol)
        shutil.copy(mdprop_path, os.path.join(d, "MDPROP"))
        p = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=timeout)
        log = p.stdout + p.stderr
        if p.returncode != 0 or not os.path.exists(os.path.join(d, "ccout.bin")):
            raise RuntimeError(f"reference harness failed (rc={p.returncode}):\n{log[-4000:]}")
        raw = np.fromfile(os.path.join(d, "ccout.bin"), dtype=np.float64)
    ndimx = int(round(raw[-2]))
    out, pos = {}, 0
    for name, n in (("T1", nt1), ("T2", nt2), ("DOO", nfoo), ("DVV", nfvv),
                    ("XVO", ndimx), ("DVO", ndimx), ("DOV", ndimx)):
        out[name] = raw[pos:pos + 2 * n].view(np.complex128).copy()
        pos += 2 * n
    out["DG"] = raw[pos:pos + ndimx].copy()
    pos += ndimx
    out["expval"] = complex(raw[pos], raw[pos + 1])
    out["cpu_s"] = float(raw[-1])
    out["log"] = log
    return out
