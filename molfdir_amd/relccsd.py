"""Host-side mirror of relccsd's CCSD amplitude path over the HIP C ABI.

``RelCCSD`` plays the part of relccsd's energy driver CCENER
(ccdriv.F:1574-1861): it takes the Fock blocks and the six sorted integral
classes in relccsd's layouts (see molfdir_amd.synth), forms the MP2 guess
(MP2EQN cceqns.F:33) and runs the CCEQN iteration (cceqns.F:141-643) -- all
on the GPU through librelccsd_hip.so.  The per-term methods carry the names
of the relccsd routines they replace.

There is no CPU fallback: constructing a solver without the built library
raises.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librelccsd_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "relccsd_hip.h")

CLASSES = {"OOOO": 0, "VOOO": 1, "VVOO": 2, "VOVO": 3, "VOVV": 4, "VVVV": 5}
ARRAYS = {"HO": 0, "HV": 1, "HOV": 2, "GO": 3, "GV": 4, "AZ": 5, "H": 6, "S1": 7,
          "S2": 8, "T1": 9, "T2": 10, "TAU": 11, "FOO": 12, "FVO": 13, "FVV": 14, "OOOO": 15,
          "VOOO": 16, "VVOO": 17, "VOVO": 18, "VOVV": 19, "VVVV": 20, "DOO": 21, "DVV": 22,
          "XVO": 23, "DVO": 24, "DOV": 25, "MP2T1": 26, "MP2T2": 27}
TERMS = ("hointm", "hvintm", "hovintm", "t1eqns", "gointm", "gvintm", "aintm", "hintm",
         "t2eqns", "bintm", "denom", "zcore", "accept")

_lib = None


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def mccres_write(path, t1, t2, imstat=None, energies=None):
    """relccsd restart file (PUTAMPT / CCDUMP layout) from host amplitudes in
    relccsd's layout; imstat 30 ints, energies ETOT ESCF EMP2 ECCSD ET1 ET2 ET3.
    No GPU needed."""
    lib = load()
    t1 = np.ascontiguousarray(t1, dtype=np.complex128)
    t2 = np.ascontiguousarray(t2, dtype=np.complex128)
    st = None if imstat is None else (ctypes.c_int * 30)(*[int(x) for x in imstat])
    en = None if energies is None else np.ascontiguousarray(energies, dtype=np.float64)
    if lib.rcc_mccres_write(str(path).encode(), t1.size, t2.size, _dptr(t1.view(np.float64)),
                            _dptr(t2.view(np.float64)), st,
                            None if en is None else _dptr(en)) != 0:
        raise RccError(lib.rcc_file_error().decode())


def mrconee_write(path, eps, irreps, nelec, h1, multb=None, ecore=0.0, breit=False):
    """MRCONEE records 1-6 (RDSYMI / GETFM layout) through the library's
    writer: irreps 1-based per spinor, nelec per irrep, h1 NSP x NSP complex,
    multb the 2NREP x 2NREP table (None: C1).  No GPU needed."""
    lib = load()
    eps = np.ascontiguousarray(eps, dtype=np.float64)
    nsp = eps.size
    irr = np.ascontiguousarray(irreps, dtype=np.int32)
    ne = np.ascontiguousarray(nelec, dtype=np.int32)
    h = np.asfortranarray(np.asarray(h1, dtype=np.complex128).reshape(nsp, nsp))
    mb = None if multb is None else np.asfortranarray(multb, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    if lib.rcc_mrconee_write(str(path).encode(), nsp, ne.size, irr.ctypes.data_as(ip),
                             ne.ctypes.data_as(ip), None if mb is None else mb.ctypes.data_as(ip),
                             _dptr(eps), _dptr(h.reshape(-1, order="F").view(np.float64)),
                             float(ecore), 1 if breit else 0) != 0:
        raise RccError(lib.rcc_file_error().decode())


def mccres_read(path, nt1, nt2):
    """(T1, T2, IMSTAT, energies) of a relccsd restart file (GETAMPT layout)."""
    lib = load()
    t1 = np.empty(nt1, np.complex128)
    t2 = np.empty(nt2, np.complex128)
    st = (ctypes.c_int * 30)()
    en = np.empty(7, np.float64)
    if lib.rcc_mccres_read(str(path).encode(), nt1, nt2, _dptr(t1.view(np.float64)),
                           _dptr(t2.view(np.float64)), st, _dptr(en)) != 0:
        raise RccError(lib.rcc_file_error().decode())
    return t1, t2, np.array(list(st)), en


class RccError(RuntimeError):
    pass


class _Sym(ctypes.Structure):
    _fields_ = [("nrep", ctypes.c_int), ("no", ctypes.POINTER(ctypes.c_int)),
                ("nv", ctypes.POINTER(ctypes.c_int)), ("frozen", ctypes.POINTER(ctypes.c_int)),
                ("multb", ctypes.POINTER(ctypes.c_int))]

# layout ids of rcc_layout_block
LAYOUTS = {"OOOO": 0, "VOOO": 1, "VVOO": 2, "VOVO": 3, "VOVV": 4, "VVVV": 5,
           "T1": 6, "T2": 7, "FOO": 8, "FVV": 9}
# layout of each readable array
ARRAY_LAYOUT = {"HO": "FOO", "GO": "FOO", "HV": "FVV", "GV": "FVV", "HOV": "T1", "S1": "T1",
                "T1": "T1", "AZ": "OOOO", "H": "VOVO", "S2": "T2", "T2": "T2", "TAU": "T2",
                "FOO": "FOO", "FVO": "T1", "FVV": "FVV", "OOOO": "OOOO", "VOOO": "VOOO",
                "VVOO": "VVOO", "VOVO": "VOVO", "VOVV": "VOVV", "VVVV": "VVVV", "DOO": "FOO",
                "DVV": "FVV", "XVO": "T1", "DVO": "T1", "DOV": "T1", "MP2T1": "T1",
                "MP2T2": "T2"}


def xor_multb(nrep: int) -> np.ndarray:
    """Pair-irrep table (1-based, column-major) of the abelian double group
    Z2^k x {fermion, boson} with nrep = 2^k fermion irreps."""
    g = np.arange(nrep)
    return (np.bitwise_xor.outer(g, g) + 1).astype(np.int32)


def declared_symbols() -> list[str]:
    """Names of every function include/relccsd_hip.h declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|double|int64_t|const char\*)\s+(rcc_\w+)\s*\(",
                                 text, re.M)))


def load():
    """Loads librelccsd_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RccError(f"{LIB_PATH} is missing: build it with `python -m molfdir_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    vp, i64, ci, dp = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    lib.rcc_create.argtypes = [ctypes.POINTER(_Sym), ctypes.POINTER(vp)]
    lib.rcc_destroy.argtypes = [vp]
    lib.rcc_destroy.restype = None
    lib.rcc_last_error.argtypes = [vp]
    lib.rcc_last_error.restype = ctypes.c_char_p
    lib.rcc_set_options.argtypes = [vp, ci, ci, ci, ci]
    lib.rcc_set_fock.argtypes = [vp, vp, vp, vp, vp]
    lib.rcc_set_complex_product.argtypes = [vp, ci]
    lib.rcc_set_ints.argtypes = [vp, ci, ci, i64, i64, vp]
    lib.rcc_set_ints_device.argtypes = [vp, ci, ci, i64, i64, vp]
    lib.rcc_set_amplitudes.argtypes = [vp, vp, vp]
    lib.rcc_get_amplitudes.argtypes = [vp, vp, vp]
    lib.rcc_get_array.argtypes = [vp, ci, vp, i64]
    lib.rcc_mp2.argtypes = [vp, dp]
    for t in TERMS:
        getattr(lib, "rcc_" + t).argtypes = [vp]
    lib.rcc_diis.argtypes = [vp, dp]
    lib.rcc_energy.argtypes = [vp, dp]
    lib.rcc_t1diag.argtypes = [vp, dp]
    lib.rcc_iterate.argtypes = [vp, ctypes.POINTER(ci), dp, dp]
    lib.rcc_cceqn.argtypes = [vp, ctypes.POINTER(ci), dp, dp]
    lib.rcc_last_iteration_ms.argtypes = [vp, dp]
    lib.rcc_kernel_times.argtypes = [vp, dp]
    lib.rcc_zgemm_device.argtypes = [vp, ci, ci, ci, ci, ci, dp, vp, i64, vp, i64, dp, vp, i64]
    lib.rcc_zgemm_device_noalloc.argtypes = [vp, ci, ci, ci, ci, ci, dp, vp, i64, vp, i64, dp,
                                             vp, i64, ci]
    lib.rcc_probe_mfma.argtypes = [vp, dp]
    lib.rcc_set_graph.argtypes = [vp, ci]
    lib.rcc_iterate_batch.argtypes = [ctypes.POINTER(vp), ci, ctypes.POINTER(ci), dp, dp]
    lib.rcc_iterate_batch_grouped.argtypes = [ctypes.POINTER(vp), ci, ctypes.POINTER(ci), dp, dp]
    lib.rcc_batch_plan_info.argtypes = [ctypes.POINTER(vp), ci, ctypes.POINTER(ci)]
    lib.rcc_layout_block.argtypes = [vp, ci, ci, ctypes.POINTER(i64), ctypes.POINTER(i64),
                                     ctypes.POINTER(i64)]
    lib.rcc_zgemm_split_device.argtypes = [vp, ci, ci, ci, dp, vp, i64, vp, i64, dp, vp, i64]
    lib.rcc_zgemm_i8_device.argtypes = [vp, ci, ci, ci, ci, dp, vp, i64, vp, i64, dp, vp, i64]
    lib.rcc_set_split_digits.argtypes = [vp, ci]
    lib.rcc_probe_mfma_rate.argtypes = [vp, ci, ci, ci, dp]
    lib.rcc_synchronize.argtypes = [vp]
    lib.rcc_lambda.argtypes = [vp, ctypes.POINTER(ci), dp]
    lib.rcc_set_autotune.argtypes = [vp, ci]
    lib.rcc_autotune_entries.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
    lib.rcc_set_lambda.argtypes = [vp, vp, vp]
    lib.rcc_get_lambda.argtypes = [vp, vp, vp]
    lib.rcc_lambda_vjp.argtypes = [vp, vp, vp, ci, vp, vp]
    lib.rcc_density.argtypes = [vp, vp]
    lib.rcc_last_lambda_ms.argtypes = [vp, dp]
    lib.rcc_mccres_write.argtypes = [ctypes.c_char_p, i64, i64, dp, dp, ctypes.POINTER(ci), dp]
    lib.rcc_mccres_read.argtypes = [ctypes.c_char_p, i64, i64, dp, dp, ctypes.POINTER(ci), dp]
    lib.rcc_file_error.restype = ctypes.c_char_p
    lib.rcc_mrconee_write.argtypes = [ctypes.c_char_p, ci, ci, ctypes.POINTER(ci),
                                      ctypes.POINTER(ci), ctypes.POINTER(ci), dp, dp,
                                      ctypes.c_double, ci]
    lib.rcc_write_mccres.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ci), dp]
    lib.rcc_read_mccres.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ci), dp]
    lib.rcc_read_molfdir.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p,
                                     ctypes.POINTER(ci), ci, dp]
    lib.rcc_read_molfdir_opts.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p,
                                          ctypes.POINTER(ci), ci, ci, ci, dp]
    lib.rcc_sort_ints_device.argtypes = [vp, ctypes.POINTER(ci), ci, ctypes.POINTER(ci), ci, vp,
                                         vp, i64, dp]
    lib.rcc_last_sort_ms.argtypes = [vp]
    lib.rcc_last_sort_ms.restype = ctypes.c_double
    lib.rcc_get_eps.argtypes = [vp, dp]
    lib.rcc_read_property.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                      ctypes.POINTER(ci), ci, vp]
    lib.rcc_expectation.argtypes = [vp, vp, ci, dp]
    lib.rcc_mp2_density.argtypes = [vp, ci, ci, ci, ctypes.POINTER(ci), dp]
    lib.rcc_last_mp2d_ms.argtypes = [vp]
    lib.rcc_last_mp2d_ms.restype = ctypes.c_double
    lib.rcc_mp2d_timing.argtypes = [vp, dp]
    lib.rcc_t3corr.argtypes = [vp, dp]
    lib.rcc_t3corr_part.argtypes = [vp, i64, i64, dp]
    lib.rcc_t3_triples.argtypes = [vp]
    lib.rcc_t3_triples.restype = i64
    lib.rcc_last_t3_ms.argtypes = [vp, dp]
    lib.rcc_last_t3_ms.restype = ctypes.c_double
    _lib = lib
    return lib


def _host(x) -> np.ndarray:
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(x))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class RelCCSD:
    """One relativistic CCSD problem resident on the current GPU.

    sym = (occupied per irrep, virtual per irrep) selects an abelian double
    group with that many fermion irreps (pair-irrep table `multb`, default the
    Z2^k table); every array of the interface is then in relccsd's
    symmetry-blocked layout, and `no` / `nv` are the totals."""

    def __init__(self, no: int, nv: int, frozen=None, maxdim: int = 8, maxit: int = 30,
                 ntol: int = 12, precision: int = 0, sym=None, multb=None):
        self.lib = load()
        self.no, self.nv = int(no), int(nv)
        self.noot, self.nvvt = self.no * (self.no - 1) // 2, self.nv * (self.nv - 1) // 2
        self.nt1, self.nt2 = self.no * self.nv, self.noot * self.nvvt
        no_l, nv_l = ([self.no], [self.nv]) if sym is None else (list(sym[0]), list(sym[1]))
        if sum(no_l) != self.no or sum(nv_l) != self.nv or len(no_l) != len(nv_l):
            raise RccError("sym counts must sum to no, nv")
        self.nrep = len(no_l)
        self._no = (ctypes.c_int * self.nrep)(*no_l)
        self._nv = (ctypes.c_int * self.nrep)(*nv_l)
        fr = None
        if frozen is not None:
            self._fr = (ctypes.c_int * max(1, self.no))(*[int(bool(x)) for x in frozen])
            fr = ctypes.cast(self._fr, ctypes.POINTER(ctypes.c_int))
        mt = None
        if self.nrep > 1:
            tab = xor_multb(self.nrep) if multb is None else np.asarray(multb, dtype=np.int32)
            self._mt = (ctypes.c_int * (self.nrep ** 2))(*tab.T.reshape(-1).tolist())
            mt = ctypes.cast(self._mt, ctypes.POINTER(ctypes.c_int))
        sym = _Sym(self.nrep, ctypes.cast(self._no, ctypes.POINTER(ctypes.c_int)),
                   ctypes.cast(self._nv, ctypes.POINTER(ctypes.c_int)), fr, mt)
        h = ctypes.c_void_p()
        rc = self.lib.rcc_create(ctypes.byref(sym), ctypes.byref(h))
        self.h = h
        if rc != 0:
            msg = self.lib.rcc_last_error(h).decode() if h.value else "rcc_create failed"
            self.close()
            raise RccError(msg)
        self.set_options(maxdim, maxit, ntol, precision)
        if self.nrep > 1:
            self.nt1, self.nt2 = self.layout_size("T1"), self.layout_size("T2")

    # ------------------------------------------------------------ layouts
    def layout_block(self, name, block):
        """(rows, cols, offset) of block `block` of a layout (SYMTAB offsets)."""
        r, c, o = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.rcc_layout_block(self.h, LAYOUTS[name], block, ctypes.byref(r),
                                              ctypes.byref(c), ctypes.byref(o)))
        return r.value, c.value, o.value

    def layout_size(self, name):
        return self.layout_block(name, self.nrep)[2]

    # ------------------------------------------------------------ plumbing
    def _check(self, rc):
        if rc != 0:
            raise RccError(self.lib.rcc_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.rcc_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_options(self, maxdim=8, maxit=30, ntol=12, precision=0):
        """precision 0: FP64; 1: split-BF16, 2: split-INT8 products for the ladder,
        VOVV.tau and the rings."""
        self.maxdim, self.maxit, self.ntol, self.precision = maxdim, maxit, ntol, precision
        self._check(self.lib.rcc_set_options(self.h, maxdim, maxit, ntol, precision))

    def set_complex_product(self, three_mult: bool):
        """Complex GEMM products: Gauss 3 real MFMA products (True, default) or 4."""
        self._check(self.lib.rcc_set_complex_product(self.h, 1 if three_mult else 0))

    # ------------------------------------------------------------ inputs
    def set_fock(self, foo, fvo, fvv, eps):
        """FMTOFILE's Fock blocks (ccgetv.F:34) and the orbital energies."""
        a = [np.ascontiguousarray(_host(x), dtype=np.complex128) for x in (foo, fvo, fvv)]
        e = np.ascontiguousarray(_host(eps), dtype=np.float64)
        self._check(self.lib.rcc_set_fock(self.h, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]), _ptr(e)))

    def set_ints(self, name, data, col0=0, ncols=None, irrep=None):
        """A column batch of one class block (irrep: compound irrep, 0-based);
        with symmetry and irrep=None, the whole class in symmetry layout."""
        cls = CLASSES[name]
        if self.nrep > 1:
            if irrep is None:
                flat = data.reshape(-1) if hasattr(data, "reshape") else data
                for b in range(self.nrep):
                    r, c, o = self.layout_block(name, b)
                    if r * c:
                        self.set_ints(name, flat[o:o + r * c], 0, c, irrep=b)
                return
            r, c, _ = self.layout_block(name, irrep)
            rows = r
        else:
            if irrep not in (None, 0):
                raise RccError("C1 problem: irrep must be 0")
            irrep = 0
        if self.nrep == 1:
            rows = {"OOOO": self.noot, "VOOO": self.nt1, "VVOO": self.nvvt, "VOVO": self.nt1,
                    "VOVV": self.nt1, "VVVV": self.nvvt}[name]
        if hasattr(data, "is_cuda") and data.is_cuda:
            import torch
            torch.cuda.synchronize()     # the producer ran on torch's stream
            n = data.numel()
            ncols = n // max(rows, 1) if ncols is None else ncols
            if n != rows * ncols or data.dtype.__str__() != "torch.complex128" or not data.is_contiguous():
                raise RccError(f"{name}: need a contiguous complex128 tensor of {rows}x{ncols}")
            self._check(self.lib.rcc_set_ints_device(self.h, cls, irrep, col0, ncols,
                                                     ctypes.c_void_p(data.data_ptr())))
            return
        a = np.ascontiguousarray(_host(data), dtype=np.complex128).reshape(-1)
        ncols = a.size // max(rows, 1) if ncols is None else ncols
        if a.size != rows * ncols:
            raise RccError(f"{name}: {a.size} elements is not {rows} x {ncols}")
        self._check(self.lib.rcc_set_ints(self.h, cls, irrep, col0, ncols, _ptr(a)))

    def set_system(self, sysd):
        """Uploads a system dict in the handle's relccsd layout (C1 arrays from
        molfdir_amd.synth, or symmetry-blocked arrays for a symmetric handle)."""
        self.set_fock(sysd["FOO"], sysd["FVO"], sysd["FVV"], sysd["eps"])
        for name in CLASSES:
            self.set_ints(name, sysd[name])

    def set_amplitudes(self, t1, t2):
        a = np.ascontiguousarray(_host(t1), dtype=np.complex128)
        b = np.ascontiguousarray(_host(t2), dtype=np.complex128)
        assert a.size == self.nt1 and b.size == self.nt2
        self._check(self.lib.rcc_set_amplitudes(self.h, _ptr(a), _ptr(b)))

    def get_amplitudes(self):
        a = np.empty(self.nt1, np.complex128)
        b = np.empty(self.nt2, np.complex128)
        self._check(self.lib.rcc_get_amplitudes(self.h, _ptr(a), _ptr(b)))
        return a, b

    def get(self, name):
        if self.nrep > 1:
            n = self.layout_size(ARRAY_LAYOUT[name])
        else:
            o, v, oo, vv, vo = self.no, self.nv, self.noot, self.nvvt, self.nt1
            n = {"HO": o ** 2, "HV": v ** 2, "HOV": vo, "GO": o ** 2, "GV": v ** 2, "AZ": oo ** 2,
                 "H": vo ** 2, "S1": vo, "S2": self.nt2, "T1": vo, "T2": self.nt2,
                 "TAU": self.nt2, "FOO": o ** 2, "FVO": vo, "FVV": v ** 2, "OOOO": oo ** 2,
                 "VOOO": vo * oo, "VVOO": vv * oo, "VOVO": vo ** 2, "VOVV": vo * vv,
                 "VVVV": vv ** 2, "DOO": o ** 2, "DVV": v ** 2, "XVO": vo, "DVO": vo,
                 "DOV": vo, "MP2T1": vo, "MP2T2": self.nt2}[name]
        out = np.empty(n, np.complex128)
        self._check(self.lib.rcc_get_array(self.h, ARRAYS[name], _ptr(out), n))
        return out

    # ------------------------------------------------------------ relccsd routines
    def mp2(self):
        """MP2EQN: T = (FVO, VVOO)/D; returns the MP2 correlation energy."""
        e = ctypes.c_double()
        self._check(self.lib.rcc_mp2(self.h, ctypes.byref(e)))
        return e.value

    def term(self, name):
        self._check(getattr(self.lib, "rcc_" + name)(self.h))

    def diis(self):
        r = ctypes.c_double()
        self._check(self.lib.rcc_diis(self.h, ctypes.byref(r)))
        return r.value

    def energy(self):
        e = ctypes.c_double()
        self._check(self.lib.rcc_energy(self.h, ctypes.byref(e)))
        return e.value

    def t1diag(self):
        d = ctypes.c_double()
        self._check(self.lib.rcc_t1diag(self.h, ctypes.byref(d)))
        return d.value

    def iterate(self):
        """One CCEQN iteration; returns (nit, Ec, rms)."""
        n, e, r = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.rcc_iterate(self.h, ctypes.byref(n), ctypes.byref(e), ctypes.byref(r)))
        return n.value, e.value, r.value

    def set_graph(self, enable: bool = True):
        """Replay the per-iteration term sequence from a captured hipGraph."""
        self._check(self.lib.rcc_set_graph(self.h, 1 if enable else 0))

    @staticmethod
    def iterate_batch(solvers, grouped=False):
        """One CCEQN iteration of every solver, their term phases in flight
        together (grouped: the four dominant products as one launch per term
        over all solvers); returns a list of (nit, Ec, rms)."""
        n = len(solvers)
        if n == 0:
            return []
        lib = solvers[0].lib
        hs = (ctypes.c_void_p * n)(*[s.h.value for s in solvers])
        nit = (ctypes.c_int * n)()
        e = (ctypes.c_double * n)()
        r = (ctypes.c_double * n)()
        fn = lib.rcc_iterate_batch_grouped if grouped else lib.rcc_iterate_batch
        rc = fn(hs, n, nit, e, r)
        if rc != 0:
            msgs = [lib.rcc_last_error(s.h).decode() for s in solvers]
            raise RccError("; ".join(m for m in msgs if m) or "rcc_iterate_batch failed")
        return [(nit[i], e[i], r[i]) for i in range(n)]

    @staticmethod
    def batch_plan_info(solvers):
        """The grouped batch plan of these solvers: compiled, launch positions,
        grouped products, and their kernels (96 x 64, 32 x 32, 32 x 32
        persistent and 128 x 32 tiles; the streaming kernel)."""
        n = len(solvers)
        lib = solvers[0].lib
        hs = (ctypes.c_void_p * n)(*[s.h.value for s in solvers])
        info = (ctypes.c_int * 8)()
        if lib.rcc_batch_plan_info(hs, n, info) != 0:
            raise RccError(lib.rcc_last_error(solvers[0].h).decode())
        return {"compiled": bool(info[0]), "launch_positions": info[1],
                "grouped_products": info[2],
                "tile_variants": {"96x64": info[3], "32x32": info[4], "32x32_persistent": info[5],
                                  "streaming": info[6], "128x32": info[7]}}

    def cceqn(self):
        """MP2 guess then the CCEQN loop until rms < 10^-ntol or maxit.
        Returns dict(emp2, rows=[(nit, Ec, rms)], t1diag=[T1DIAG of every
        iteration, as CCEQN prints it beside Ec], eccsd, nit)."""
        emp2 = self.mp2()
        rows, diag = [], []
        conv = 10.0 ** (-self.ntol)
        for _ in range(self.maxit):
            nit, e, r = self.iterate()
            rows.append((nit, e, r))
            diag.append(self.t1diag())
            if r < conv:
                break
        return {"emp2": emp2, "rows": rows, "t1diag": diag,
                "eccsd": rows[-1][1] if rows else 0.0, "nit": len(rows)}

    def last_iteration_ms(self):
        ms = ctypes.c_double()
        self._check(self.lib.rcc_last_iteration_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def set_autotune(self, enable: bool = True):
        """Process-wide GEMM autotuner (K-split of FP64 products, INT8 tile height)."""
        self._check(self.lib.rcc_set_autotune(self.h, 1 if enable else 0))

    def autotune_entries(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.rcc_autotune_entries(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    # ------------------------------------------------------------ Lambda / density
    def lambda_solve(self):
        """Solve the CCSD Lambda equations at the current T (after cceqn);
        returns (nit, rms)."""
        n, r = ctypes.c_int(), ctypes.c_double()
        self._check(self.lib.rcc_lambda(self.h, ctypes.byref(n), ctypes.byref(r)))
        return n.value, r.value

    def set_lambda(self, l1, l2):
        a = np.ascontiguousarray(_host(l1), dtype=np.complex128)
        b = np.ascontiguousarray(_host(l2), dtype=np.complex128)
        assert a.size == self.nt1 and b.size == self.nt2
        self._check(self.lib.rcc_set_lambda(self.h, _ptr(a), _ptr(b)))

    def get_lambda(self):
        a = np.empty(self.nt1, np.complex128)
        b = np.empty(self.nt2, np.complex128)
        self._check(self.lib.rcc_get_lambda(self.h, _ptr(a), _ptr(b)))
        return a, b

    def lambda_vjp(self, l1, l2, with_energy=False):
        """with_energy * dE/dT + Lam . dS/dT at the current T (S before DENOM)."""
        a = np.ascontiguousarray(_host(l1), dtype=np.complex128)
        b = np.ascontiguousarray(_host(l2), dtype=np.complex128)
        assert a.size == self.nt1 and b.size == self.nt2
        g1 = np.empty(self.nt1, np.complex128)
        g2 = np.empty(self.nt2, np.complex128)
        self._check(self.lib.rcc_lambda_vjp(self.h, _ptr(a), _ptr(b), 1 if with_energy else 0,
                                            _ptr(g1), _ptr(g2)))
        return g1, g2

    def density(self):
        """Hermitian response density D (n x n, occupied then virtual):
        d Re(Ec)/d lam = tr(D V) for f -> f + lam V."""
        n = self.no + self.nv
        d = np.empty((n, n), np.complex128, order="F")
        self._check(self.lib.rcc_density(self.h, _ptr(d)))
        return d

    def last_lambda_ms(self):
        ms = ctypes.c_double()
        self._check(self.lib.rcc_last_lambda_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def read_molfdir(self, directory, nelec, mrconee="MRCONEE", mdcint="MDCINT", useoe=False,
                     norecmp=False):
        """Integral front end (RDINTS): the six classes, the Fock blocks and
        the orbital energies from the MOLFDIR files in `directory`; nelec the
        occupied spinors per irrep of MRCONEE; useoe / norecmp namelist
        CCSORT's options.  Returns E0, E1, ECORE and ESCF = E0 + E1 + ECORE."""
        import os
        ne = (ctypes.c_int * len(nelec))(*[int(x) for x in nelec])
        out = (ctypes.c_double * 3)()
        self._check(self.lib.rcc_read_molfdir_opts(
            self.h, os.path.join(directory, mrconee).encode(), os.path.join(directory, mdcint).encode(),
            ne, len(nelec), int(bool(useoe)), int(bool(norecmp)), out))
        return {"e0": out[0], "e1": out[1], "ecore": out[2], "escf": out[0] + out[1] + out[2]}

    def sort_ints_device(self, spinor_pos, kr, idx, val):
        """Expansion + sort of a canonical list in device memory (torch
        tensors: idx int32 (m, 4), val complex128 (m,)); returns device ms."""
        sp = np.ascontiguousarray(spinor_pos, dtype=np.int32)
        k = np.ascontiguousarray(kr, dtype=np.int32).reshape(-1)
        ms = ctypes.c_double()
        self._check(self.lib.rcc_sort_ints_device(
            self.h, sp.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), sp.size,
            k.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), k.size // 2,
            ctypes.c_void_p(idx.data_ptr()), ctypes.c_void_p(val.data_ptr()), int(idx.shape[0]),
            ctypes.byref(ms)))
        return ms.value

    def last_sort_ms(self):
        return float(self.lib.rcc_last_sort_ms(self.h))

    def read_property(self, mrconee, mdprop, name, nelec):
        """RDPROP: operator `name` of an MDPROP file as an (no+nv) x (no+nv)
        complex matrix in the handle's spinor order."""
        n = self.no + self.nv
        out = np.empty((n, n), np.complex128, order="F")
        ne = (ctypes.c_int * len(nelec))(*[int(x) for x in nelec])
        self._check(self.lib.rcc_read_property(self.h, str(mrconee).encode(), str(mdprop).encode(),
                                               name.encode(), ne, len(nelec), _ptr(out)))
        return out

    def mp2_density(self, maxit=30, maxdim=8, ntol=12):
        """CCFOPR's MP2 relaxed density (MP2EQN 'LAMBDA', CCDENSC(1), CCDENSZ(1)
        with the RPAA Z-vector); returns (iterations, last RMS)."""
        nit = ctypes.c_int()
        rms = ctypes.c_double()
        self._check(self.lib.rcc_mp2_density(self.h, int(maxit), int(maxdim), int(ntol),
                                             ctypes.byref(nit), ctypes.byref(rms)))
        return nit.value, rms.value

    def last_mp2d_ms(self):
        """Device ms of the last mp2_density: total and its phases."""
        v = (ctypes.c_double * 6)()
        self._check(self.lib.rcc_mp2d_timing(self.h, v))
        return {"total": v[0], "densities": v[1], "x": v[2], "zvector": v[3], "zgemv": v[4],
                "xgemm": v[5]}

    def expectation(self, prop, level=1):
        """EXPVAL: level 0 reference determinant, 1 unrelaxed CCSD (Lambda),
        2 MP2 relaxed (after mp2_density)."""
        p = np.asfortranarray(prop, dtype=np.complex128)
        v = (ctypes.c_double * 2)()
        self._check(self.lib.rcc_expectation(self.h, _ptr(p), int(level), v))
        return complex(v[0], v[1])

    def get_eps(self):
        out = np.empty(self.no + self.nv, np.float64)
        self._check(self.lib.rcc_get_eps(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def write_mccres(self, path, imstat=None, energies=None):
        """The handle's T1, T2 as a relccsd restart file (PUTAMPT + CCDUMP)."""
        st = None if imstat is None else (ctypes.c_int * 30)(*[int(x) for x in imstat])
        en = None if energies is None else np.ascontiguousarray(energies, dtype=np.float64)
        self._check(self.lib.rcc_write_mccres(self.h, str(path).encode(), st,
                                              None if en is None else _dptr(en)))

    def read_mccres(self, path):
        """Sets T1, T2 from a relccsd restart file (GETAMPT); returns
        (IMSTAT, energies)."""
        st = (ctypes.c_int * 30)()
        en = np.empty(7, np.float64)
        self._check(self.lib.rcc_read_mccres(self.h, str(path).encode(), st, _dptr(en)))
        return np.array(list(st)), en

    def t3corr(self, first=0, count=-1):
        """Non-iterative triples of T3CORR on the current amplitudes:
        {'et1': +T (4th order), 'et2': (T) 5th-order term, 'et3': -T 5th-order
        term}; CCSD(T) = Ec + et1 + et2.  first / count select a range of the
        occupied triples i > j > k (reference loop order)."""
        et = (ctypes.c_double * 3)()
        if first == 0 and count < 0:
            self._check(self.lib.rcc_t3corr(self.h, et))
        else:
            self._check(self.lib.rcc_t3corr_part(self.h, int(first), int(count), et))
        return {"et1": et[0], "et2": et[1], "et3": et[2]}

    def t3_triples(self):
        return int(self.lib.rcc_t3_triples(self.h))

    def last_t3_ms(self):
        """(total ms, {'gemm': ms, 'energy': ms, 'chunk': triples per launch,
        'launches': n}) of the last t3corr call."""
        p = (ctypes.c_double * 4)()
        tot = self.lib.rcc_last_t3_ms(self.h, p)
        return tot, {"gemm": p[0], "energy": p[1], "chunk": int(p[2]), "launches": int(p[3])}

    def kernel_times(self):
        """HIP-event ms of the last iteration's ladder, two rings and VOVV.tau."""
        out = (ctypes.c_double * 4)()
        self._check(self.lib.rcc_kernel_times(self.h, out))
        return {"ladder": out[0], "ring_hintm": out[1], "ring_t2": out[2], "vovv_tau": out[3]}

    def synchronize(self):
        self._check(self.lib.rcc_synchronize(self.h))

    # ------------------------------------------------------------ diagnostics
    def zgemm_device(self, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
        al = (ctypes.c_double * 2)(alpha.real, alpha.imag)
        be = (ctypes.c_double * 2)(beta.real, beta.imag)
        self._check(self.lib.rcc_zgemm_device(
            self.h, opa, opb, m, n, k, al, ctypes.c_void_p(A.data_ptr()), lda,
            ctypes.c_void_p(B.data_ptr()), ldb, be, ctypes.c_void_p(C.data_ptr()), ldc))

    def zgemm_device_noalloc(self, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                             ksplit=0):
        """rcc_zgemm_device without allocation: enqueued on the handle's
        stream (call synchronize()), K-split partial sums in the library's
        fixed 32 MiB scratch; ksplit 0 = heuristic split."""
        al = (ctypes.c_double * 2)(alpha.real, alpha.imag)
        be = (ctypes.c_double * 2)(beta.real, beta.imag)
        self._check(self.lib.rcc_zgemm_device_noalloc(
            self.h, opa, opb, m, n, k, al, ctypes.c_void_p(A.data_ptr()), lda,
            ctypes.c_void_p(B.data_ptr()), ldb, be, ctypes.c_void_p(C.data_ptr()), ldc,
            int(ksplit)))

    def mfma_rate(self, chains=8, iters=20000, nblocks=2048):
        """Measured FP64 MFMA throughput (TFLOP/s) of a register-only loop."""
        t = ctypes.c_double()
        self._check(self.lib.rcc_probe_mfma_rate(self.h, chains, iters, nblocks, ctypes.byref(t)))
        return t.value

    def zgemm_split_device(self, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
        al = (ctypes.c_double * 2)(alpha.real, alpha.imag)
        be = (ctypes.c_double * 2)(beta.real, beta.imag)
        self._check(self.lib.rcc_zgemm_split_device(
            self.h, m, n, k, al, ctypes.c_void_p(A.data_ptr()), lda,
            ctypes.c_void_p(B.data_ptr()), ldb, be, ctypes.c_void_p(C.data_ptr()), ldc))

    def zgemm_i8_device(self, digits, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
        al = (ctypes.c_double * 2)(alpha.real, alpha.imag)
        be = (ctypes.c_double * 2)(beta.real, beta.imag)
        self._check(self.lib.rcc_zgemm_i8_device(
            self.h, digits, m, n, k, al, ctypes.c_void_p(A.data_ptr()), lda,
            ctypes.c_void_p(B.data_ptr()), ldb, be, ctypes.c_void_p(C.data_ptr()), ldc))

    def set_split_digits(self, digits: int):
        """Digits (4 or 5) per value of the split-INT8 mode (precision 2)."""
        self._check(self.lib.rcc_set_split_digits(self.h, digits))

    def probe_mfma(self):
        out = np.empty(256, np.float64)
        self._check(self.lib.rcc_probe_mfma(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out.reshape(64, 4)
