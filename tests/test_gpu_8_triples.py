"""GPU checks of the non-iterative triples (rcc_t3corr, T3CORR
cceqns.F:1336-1863) against the compiled reference's own ET1 / ET2 / ET3
(tests/golden/t3_*.npz) and the oracle (oracle/triples_oracle.py).

Bar: relative 1e-11 of max(1, |E|) for each correction (FP64, sums in a
different order than the reference's loops; the reference and the oracle
agree to 5e-14 on these fixtures), and 1e-10 Eh absolute for the physical
case on the converged amplitudes.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from molfdir_amd import synth  # noqa: E402
from molfdir_amd.relccsd import RelCCSD  # noqa: E402
from oracle import triples_oracle as to  # noqa: E402

from _golden import load, names, layout_for, system_for  # noqa: E402
from test_oracle_triples import amplitudes  # noqa: E402

CLASSES = ("OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV")
KEYS = ("et1", "et2", "et3")


def handle(g, symmetric=False, **kw):
    s = system_for(g)
    no, nv = int(g["params"][0]), int(g["params"][1])
    t1, t2 = amplitudes(g)
    if symmetric:
        lay = layout_for(g)
        r = RelCCSD(no, nv, sym=(g["sym_no"].tolist(), g["sym_nv"].tolist()), **kw)
        r.set_fock(lay.to_sym("FOO", s["FOO"].numpy()), lay.to_sym("FVO", s["FVO"].numpy()),
                   lay.to_sym("FVV", s["FVV"].numpy()), s["eps"])
        for c in CLASSES:
            r.set_ints(c, lay.to_sym(c, s[c].numpy()))
        r.set_amplitudes(lay.to_sym("T1", t1), lay.to_sym("T2", t2))
    else:
        r = RelCCSD(no, nv, **kw)
        r.set_system(s)
        r.set_amplitudes(t1, t2)
    return r, s, t1, t2


def check(e, ref, rel=1e-11):
    for k, x in zip(KEYS, ref):
        assert abs(e[k] - x) <= rel * max(1.0, abs(x)), (k, e[k], x)


def ref_of(g):
    return float(g["ET1"]), float(g["ET2"]), float(g["ET3"])


@pytest.mark.parametrize("name", names("t3"))
def test_t3corr_matches_reference(name):
    g = load("t3", name)
    print(f"t3 fixture {name}", flush=True)
    r, _, _, _ = handle(g)
    check(r.t3corr(), ref_of(g))


@pytest.mark.parametrize("name", ["s4_5o12v", "s8_8o24v"])
def test_symmetric_handle_matches_reference(name):
    g = load("t3", name)
    r, _, _, _ = handle(g, symmetric=True)
    check(r.t3corr(), ref_of(g))


def test_four_product_form_and_chunked_launches_agree():
    g = load("t3", "r_7o70v")
    r, _, _, _ = handle(g)
    whole = r.t3corr()
    _, parts = r.last_t3_ms()
    assert parts["launches"] == 1
    r.set_complex_product(False)
    check(r.t3corr(), ref_of(g))
    r.set_complex_product(True)
    # a 3 MB work budget: one occupied triple per launch (35 launches)
    os.environ["RCC_T3_BUDGET_MB"] = "3"
    try:
        chunked = r.t3corr()
        _, parts = r.last_t3_ms()
    finally:
        del os.environ["RCC_T3_BUDGET_MB"]
    assert parts["launches"] == 35 and parts["chunk"] == 1
    check(chunked, [whole[k] for k in KEYS], 1e-13)


def test_triple_ranges_add_up():
    g = load("t3", "r_6o20v")
    r, _, _, _ = handle(g)
    n = r.t3_triples()
    assert n == 20
    whole = r.t3corr()
    acc = np.zeros(3)
    for first, count in ((0, 7), (7, 1), (8, 12)):
        e = r.t3corr(first, count)
        acc += [e[k] for k in KEYS]
    check(dict(zip(KEYS, acc)), [whole[k] for k in KEYS], 1e-13)
    with pytest.raises(Exception):
        r.t3corr(15, 6)


def test_ccsd_t_on_gpu_converged_amplitudes():
    # CCSD on the GPU, then (T): the reference's ET on its own converged
    # amplitudes (same system, same convergence threshold)
    g = load("t3", "conv_6o40v")
    c = load("ccsd", "cfg1_6o40v")
    s = system_for(g)
    r = RelCCSD(6, 40, maxit=20, maxdim=6, ntol=12)
    r.set_system(s)
    out = r.cceqn()
    assert abs(out["eccsd"] - float(c["eccsd"])) < 1e-10
    e = r.t3corr()
    for k, x in zip(KEYS, ref_of(g)):
        assert abs(e[k] - x) < 1e-10, (k, e[k], x)


@pytest.mark.parametrize("no,nv", [(5, 130), (18, 8)])
def test_oracle_on_larger_random_cases(no, nv):
    # three 64-row tiles of virtuals; hole segments of two K-steps (o > 16)
    s = synth.make_system(no, nv, seed=12, scale=0.02, fock_perturb=0.02)
    rng = np.random.default_rng(3)
    r = RelCCSD(no, nv)
    r.set_system(s)
    t1 = 0.01 * (rng.normal(size=r.nt1) + 1j * rng.normal(size=r.nt1))
    t2 = 0.01 * (rng.normal(size=r.nt2) + 1j * rng.normal(size=r.nt2))
    r.set_amplitudes(t1, t2)
    e = r.t3corr()
    check(e, to.t3corr(s, t1, t2))


def test_failed_work_allocation_is_reported_again_not_launched(monkeypatch):
    """Regression: when a work buffer of the triples cannot be allocated the
    call fails, and a second call with the same sizes fails the same way
    (the partially built work state is released, never launched on).  A
    budget of 1 PB asks for 16384 triples per launch, whose U' blocks
    (16384 x 140 x 9730 x 16 B = 357 GB at 48o/140v) exceed the device."""
    from molfdir_amd.relccsd import RccError
    no, nv = 48, 140
    need = 16384 * nv * (nv * (nv - 1) // 2) * 16
    if torch.cuda.get_device_properties(0).total_memory >= need:
        pytest.skip("the device holds the oversized work buffer")
    r = RelCCSD(no, nv)
    monkeypatch.setenv("RCC_T3_BUDGET_MB", str(10 ** 9))
    for _ in range(2):
        with pytest.raises(RccError, match="hipMalloc"):
            r.t3corr()
    monkeypatch.delenv("RCC_T3_BUDGET_MB")
    g = load("t3", "r_4o10v")
    check(handle(g)[0].t3corr(), ref_of(g))
