"""The symmetry-blocked engine against the symmetric handle in the C1
working layout (RCC_BLOCKED=0: C1 arrays with the ladder, VOVV.tau and rings
grouped over the irreps) on systems larger than the reference fixtures:
Dirichlet(2) irrep proportions with 8 and 4 irreps.  Per-term intermediates
to 1e-12 relative, the MP2 energy and CCSD iterations to 1e-12 Eh, the
amplitudes to 1e-12; the blocked handle's device memory a fraction of the
C1-layout one's; Lambda and (T) (C1-layout paths of a blocked handle) equal.
"""
import gc
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from molfdir_amd import synth  # noqa: E402
from molfdir_amd.relccsd import RelCCSD  # noqa: E402

CLASSES = ("OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV")
TERMS = [("hointm", "HO"), ("hvintm", "HV"), ("hovintm", "HOV"), ("t1eqns", "S1"),
         ("gointm", "GO"), ("gvintm", "GV"), ("aintm", "AZ"), ("hintm", "H"),
         ("t2eqns", "S2"), ("bintm", "S2")]


def sym_split(no, nv, nrep, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    p = rng.dirichlet([2.0] * nrep)
    return synth.dirichlet_split(no, p), synth.dirichlet_split(nv, p)


def handle(sysd, no, nv, sym, blocked, **kw):
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    free0 = torch.cuda.mem_get_info()[0]
    os.environ["RCC_BLOCKED"] = "1" if blocked else "0"
    try:
        r = RelCCSD(no, nv, sym=sym, **kw)
    finally:
        os.environ.pop("RCC_BLOCKED", None)
    maps = synth.block_maps(sym)
    r.set_fock(*(synth.to_blocks(n, sysd[n], sym, maps).cpu() for n in ("FOO", "FVO", "FVV")),
               sysd["eps"])
    for c in CLASSES:
        r.set_ints(c, synth.to_blocks(c, sysd[c], sym, maps).contiguous())
    r.synchronize()
    torch.cuda.empty_cache()
    return r, free0 - torch.cuda.mem_get_info()[0]


def release(*hs):
    for r in hs:
        r.close()
    gc.collect()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("no,nv,nrep,seed", [(16, 120, 8, 5), (12, 96, 4, 9)])
def test_blocked_engine_equals_c1_layout_handle(no, nv, nrep, seed):
    sym = sym_split(no, nv, nrep, seed)
    print(f"blocked vs C1 layout: {no}o/{nv}v, {nrep} irreps, occupied {sym[0]}, "
          f"virtual {sym[1]}", flush=True)
    sysd = synth.make_system(no, nv, seed=seed, scale=0.02, device="cuda", sym=sym)
    b, mem_b = handle(sysd, no, nv, sym, True, maxit=6, maxdim=4, ntol=14)
    c, mem_c = handle(sysd, no, nv, sym, False, maxit=6, maxdim=4, ntol=14)
    try:
        assert mem_b < 0.6 * mem_c, (mem_b, mem_c)
        # per term on random symmetric amplitudes
        rng = np.random.Generator(np.random.PCG64(seed + 1))
        n1, n2 = b.layout_size("T1"), b.layout_size("T2")
        t1 = 0.03 * (rng.normal(size=n1) + 1j * rng.normal(size=n1))
        t2 = 0.03 * (rng.normal(size=n2) + 1j * rng.normal(size=n2))
        for r in (b, c):
            r.set_amplitudes(t1, t2)
        for term, key in TERMS:
            b.term(term)
            c.term(term)
            x, y = b.get(key), c.get(key)
            assert np.abs(x - y).max() <= 1e-12 * max(1.0, np.abs(y).max()), (term, key)
        for r in (b, c):
            r.term("denom")
            r.term("zcore")
        assert np.abs(b.get("S2") - c.get("S2")).max() < 1e-12
        # MP2 and the iterations (DIIS, energy) from the MP2 guess
        eb, ec = b.mp2(), c.mp2()
        assert abs(eb - ec) < 1e-12
        for _ in range(5):
            rb, rc = b.iterate(), c.iterate()
            assert rb[0] == rc[0] and abs(rb[1] - rc[1]) < 1e-12 and abs(rb[2] - rc[2]) < 1e-10
        for x, y in zip(b.get_amplitudes(), c.get_amplitudes()):
            assert np.abs(x - y).max() < 1e-12
    finally:
        release(b, c)


def test_c1_layout_paths_of_a_blocked_handle():
    # Lambda and (T) run in the C1 layout: a blocked handle materialises it
    # from its blocks and must agree with the C1-layout handle
    no, nv, nrep, seed = 8, 48, 4, 21
    sym = sym_split(no, nv, nrep, seed)
    sysd = synth.make_system(no, nv, seed=seed, scale=0.02, device="cuda", sym=sym)
    b, _ = handle(sysd, no, nv, sym, True, maxit=40, maxdim=6, ntol=11)
    c, _ = handle(sysd, no, nv, sym, False, maxit=40, maxdim=6, ntol=11)
    try:
        for r in (b, c):
            r.cceqn()
        eb, ec = b.t3corr(), c.t3corr()
        for k in eb:
            assert abs(eb[k] - ec[k]) <= 1e-11 * abs(ec[k]) + 1e-15, (k, eb[k], ec[k])
        lb, lc = b.lambda_solve(), c.lambda_solve()
        assert lb[0] == lc[0]
        for x, y in zip(b.get_lambda(), c.get_lambda()):
            assert np.abs(x - y).max() < 1e-10
    finally:
        release(b, c)
