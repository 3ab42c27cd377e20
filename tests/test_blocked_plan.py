"""The symmetry-blocked iteration's host-side plan (index maps, grouped
product descriptors with their K splits, the term program) run on the CPU by
a plain-loop emulator (tests/blocked_emu.cpp, test infrastructure) against
the reference's symmetric per-term fixtures (2, 4 and 8 irreps, frozen core)
and the dense oracle on larger systems whose products take the K-split path.
The GPU tests run the same plan on the device."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from _golden import (load, names, system_for, layout_for, symmetric_amplitudes,
                     frozen_mask_for, TERM_LAYOUT)

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "molfdir_amd", "csrc")
SEQ = [("HOINTM", "HO", "HO"), ("HVINTM", "HV", "HV"), ("HOVINTM", "HOV", "HOV"),
       ("T1EQNS", "S1", "S1"), ("GOINTM", "GO", "GO"), ("GVINTM", "GV", "GV"),
       ("AINTM", "AZ", "AZ"), ("HINTM", "H", "H"), ("T2EQNS", "S2", "S2_T2EQNS"),
       ("BINTM", "S2", "S2_BINTM")]


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("emu") / "blocked_emu.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-pthread", "-I" + CSRC,
                    os.path.join(HERE, "blocked_emu.cpp"), "-o", out], check=True)
    lib = ctypes.CDLL(out)
    lib.emu_create.restype = ctypes.c_void_p
    lib.emu_size.restype = ctypes.c_int64
    lib.emu_size.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.emu_arr.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    lib.emu_prog.argtypes = [ctypes.c_char_p]
    for f in ("emu_set", "emu_get"):
        getattr(lib, f).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    lib.emu_set_eps.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.emu_run.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.emu_nprods.argtypes = [ctypes.c_void_p]
    lib.emu_prod_info.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    return lib


class Blocked:
    def __init__(self, lib, no, nv, frozen=None):
        self.lib = lib
        n = len(no)
        ptab = np.array([i ^ j for j in range(n) for i in range(n)], dtype=np.int32)
        a = lambda x: np.ascontiguousarray(x, dtype=np.int32)
        self.keep = [a(no), a(nv), ptab, None if frozen is None else a(frozen)]
        ptr = lambda x: None if x is None else x.ctypes.data_as(ctypes.c_void_p)
        self.h = lib.emu_create(n, ptr(self.keep[0]), ptr(self.keep[1]), ptr(self.keep[2]),
                                ptr(self.keep[3]))
        assert self.h

    def _id(self, name):
        i = self.lib.emu_arr(self.h, name.encode())
        assert i >= 0, name
        return i

    def set(self, name, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        i = self._id(name)
        assert x.size == self.lib.emu_size(self.h, i), name
        self.lib.emu_set(self.h, i, x.ctypes.data_as(ctypes.c_void_p))

    def get(self, name):
        i = self._id(name)
        x = np.zeros(self.lib.emu_size(self.h, i), dtype=np.complex128)
        self.lib.emu_get(self.h, i, x.ctypes.data_as(ctypes.c_void_p))
        return x

    def run(self, prog):
        p = self.lib.emu_prog(prog.encode())
        assert p >= 0, prog
        self.lib.emu_run(self.h, p)

    def splits(self):
        info, n = np.zeros(5), 0
        for i in range(self.lib.emu_nprods(self.h)):
            self.lib.emu_prod_info(self.h, i, info.ctypes.data_as(ctypes.c_void_p))
            n += info[1] > 0
        return n

    def terms(self, lay, s, t1, t2, ref, tol=1e-12):
        for c in ("FOO", "FVO", "FVV", "OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV"):
            self.set(c, lay.to_sym(c, s[c].numpy()))
        eps = np.ascontiguousarray(s["eps"].numpy(), dtype=np.float64)
        self.lib.emu_set_eps(self.h, eps.ctypes.data_as(ctypes.c_void_p))
        self.run("CONST")
        self.set("T1", lay.to_sym("T1", t1))
        self.set("T2", lay.to_sym("T2", t2))
        self.run("TAU")
        self.run("RING")
        self.run("X")
        for prog, arr, key in SEQ:
            if prog == "HINTM":
                self.run("HINTM_HAKCI")
            self.run(prog)
            got, want = self.get(arr), ref(key)
            err = np.abs(got - want).max()
            assert err <= tol * max(1.0, np.abs(want).max()), (prog, err)
        self.run("DENOM")
        for arr, key in (("S1", "S1_DENOM"), ("S2", "S2_DENOM")):
            assert np.abs(self.get(arr) - ref(key)).max() < tol, key


@pytest.mark.parametrize("name", names("symterms"))
def test_blocked_plan_matches_reference_terms(emu, name):
    g = load("symterms", name)
    s, lay = system_for(g), layout_for(g)
    t1, t2 = symmetric_amplitudes(lay, int(g["params"][3]))
    Blocked(emu, g["sym_no"].tolist(), g["sym_nv"].tolist()).terms(lay, s, t1, t2,
                                                                   lambda k: g[k])


@pytest.mark.parametrize("name", [n for n in names("frzterms") if "s" in n.split("_")[0]])
def test_blocked_plan_frozen_core_matches_reference(emu, name):
    g = load("frzterms", name)
    s, lay = system_for(g), layout_for(g)
    t1, t2 = symmetric_amplitudes(lay, int(g["params"][3]))
    b = Blocked(emu, g["sym_no"].tolist(), g["sym_nv"].tolist(), frozen_mask_for(g))
    b.terms(lay, s, t1, t2, lambda k: g[k])


@pytest.mark.parametrize("sym", [([5, 4, 4, 3], [24, 20, 20, 16]),
                                 ([3, 2, 2, 1, 2, 1, 1, 0], [12, 10, 9, 8, 11, 7, 9, 6])])
def test_blocked_plan_matches_dense_oracle_with_k_splits(emu, sym):
    from oracle import ccsd_oracle
    from oracle.symtab import SymLayout
    from molfdir_amd import synth
    no_l, nv_l = sym
    s = synth.make_system(sum(no_l), sum(nv_l), seed=11, scale=0.05, fock_perturb=0.02, sym=sym)
    lay = SymLayout(no_l, nv_l)
    t1, t2 = symmetric_amplitudes(lay, 5)
    ref = ccsd_oracle.terms(s, t1, t2)
    b = Blocked(emu, no_l, nv_l)
    assert b.splits() >= 1          # long products take the K-split path
    b.terms(lay, s, t1, t2, lambda k: lay.to_sym(TERM_LAYOUT[k], np.asarray(ref[k])))
