"""MP2 relaxed first-order property on the GPU: rcc_mp2_density (CCFOPR's
DOMP2G block, ccdriv.F:2058-2090) and rcc_expectation level 2 (EXPVAL
ccgrad.F:261-347).

* Against relccsd's own chain (harness mode 7, tests/golden/mp2g_*.npz): the
  MP2 amplitudes, DOO, DVV and the X intermediate to 1e-12 relative; the
  Z-vector (DVO, DOV), converged to the same DIIS RMS as the reference, to
  1e-10; the expectation value to 1e-11.
* Against the oracle's direct solve of the Z-vector equations at 10o/60v
  (600 VO pairs): X to 1e-12, the Z-vector to 1e-10.
* Symmetry-blocked handles (2 and 4 irreps) against the reference's
  symmetric runs, in SYMTAB's layouts.
* At 20o/200v (4000 VO pairs) the Z-vector satisfies the response
  equations built from the handle's own classes, to the convergence
  threshold.
* Two calls give bit-identical densities; the CCSD amplitudes of the handle
  are left alone; level 2 or the density arrays before the density, and a
  DIIS dimension beyond 16, are refused.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from molfdir_amd import synth  # noqa: E402
from molfdir_amd.relccsd import RelCCSD, RccError  # noqa: E402
from oracle import mp2grad_oracle as O  # noqa: E402
from oracle.ccsd_oracle import Layout  # noqa: E402

from _golden import load, names, system_for  # noqa: E402


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300)


def solver_for(s, sym=None, **kw):
    """A handle holding system s: C1, or symmetry-blocked with the classes
    handed over in SYMTAB's layout."""
    r = RelCCSD(int(s["no"]), int(s["nv"]), sym=sym, **kw)
    if sym is None:
        r.set_system(s)
        return r
    from oracle.symtab import SymLayout
    lay = SymLayout(*sym)
    r.set_fock(*(lay.to_sym(k, s[k].numpy()) for k in ("FOO", "FVO", "FVV")), s["eps"])
    for name in ("OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV"):
        r.set_ints(name, lay.to_sym(name, s[name].numpy()))
    return r


@pytest.mark.parametrize("name", names("mp2g"))
def test_mp2_density_matches_reference(name):
    """C1 fixtures, and symmetry-blocked handles (2 and 4 irreps, arrays in
    SYMTAB's layouts; the reference's VO vectors are zero beyond their
    totally symmetric block)."""
    g = load("mp2g", name)
    no, nv, _, pseed, maxdim, ntol = (int(x) for x in g["params"])
    s = system_for(g)
    sym = s.get("sym")
    P = synth.hermitian_perturbation(no + nv, pseed, 0.1,
                                     irreps=None if sym is None else synth.spinor_irreps(sym))
    r = solver_for(s, sym=sym)
    print(f"mp2g {name}: no={no} nv={nv} sym={sym} maxdim={maxdim} ntol={ntol}", flush=True)
    nit, rms = r.mp2_density(maxit=100, maxdim=maxdim, ntol=ntol)
    assert rms < 10.0 ** -ntol and nit < 100
    nt1 = g["T1"].size
    assert rel(r.get("MP2T1"), g["T1"]) < 1e-13
    assert rel(r.get("DOO"), g["DOO"]) < 1e-12
    assert rel(r.get("DVV"), g["DVV"]) < 1e-12
    assert rel(r.get("XVO"), g["XVO"][:nt1]) < 1e-12
    assert rel(r.get("DVO"), g["DVO"][:nt1]) < 1e-10
    assert rel(r.get("DOV"), g["DOV"][:nt1]) < 1e-10
    e = r.expectation(P, level=2)
    assert abs(e - complex(g["expval"])) < 1e-11, (e, g["expval"])
    # the reference-determinant part is the same operator trace
    assert abs(r.expectation(P, level=0) - np.trace(P[:no, :no])) < 1e-15


def test_mp2_density_matches_oracle_10o60v():
    no, nv = 10, 60
    s = synth.make_system(no, nv, seed=11, scale=0.03, fock_perturb=0.02)
    P = synth.hermitian_perturbation(no + nv, 31, 0.1)
    o = O.mp2_relaxed(s, P)
    L = Layout(no, nv)
    r = solver_for(s)
    print(f"mp2g oracle 10o60v", flush=True)
    nit, rms = r.mp2_density(maxit=100, maxdim=8, ntol=13)
    assert rms < 1e-13
    assert rel(r.get("DOO"), L.pmat(o["doo"])) < 1e-12
    assert rel(r.get("DVV"), L.pmat(o["dvv"])) < 1e-12
    assert rel(r.get("XVO"), o["xvo"]) < 1e-12
    assert rel(r.get("DVO"), o["dvo"]) < 1e-10
    assert abs(r.expectation(P, level=2) - o["expval"]) < 1e-11


def test_repeat_is_bit_identical_and_amplitudes_untouched():
    g = load("mp2g", "g_6o20v")
    s = system_for(g)
    r = solver_for(s)
    r.mp2()
    t1, t2 = r.get("T1"), r.get("T2")
    r.mp2_density(maxit=100, maxdim=6, ntol=13)
    first = {k: r.get(k) for k in ("DOO", "DVV", "XVO", "DVO", "DOV")}
    r.mp2_density(maxit=100, maxdim=6, ntol=13)
    for k, v in first.items():
        assert np.array_equal(r.get(k), v), k
    assert np.array_equal(r.get("T1"), t1) and np.array_equal(r.get("T2"), t2)


def test_zero_right_hand_side_gives_zero_zvector():
    """No VO coupling in the X intermediate (all integrals zero but the Fock
    diagonal): X = 0, and the Z-vector stays exactly zero."""
    g = load("mp2g", "g_4o10v")
    s = system_for(g)
    for k in ("OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV", "FVO"):
        s[k] = torch.zeros_like(s[k])
    r = solver_for(s)
    nit, rms = r.mp2_density(maxit=10, maxdim=4, ntol=12)
    assert rms == 0.0 and nit == 1
    for k in ("XVO", "DVO", "DOV", "DOO", "DVV"):
        assert not np.any(r.get(k)), k
    P = synth.hermitian_perturbation(14, 3, 0.1)
    assert abs(r.expectation(P, level=2) - np.trace(P[:4, :4])) < 1e-15


def test_refusals():
    g = load("mp2g", "g_4o10v")
    s = system_for(g)
    r = solver_for(s)
    P = np.eye(14, dtype=np.complex128)
    with pytest.raises(RccError):
        r.expectation(P, level=2)
    with pytest.raises(RccError):
        r.mp2_density(maxdim=17)
    with pytest.raises(RccError):
        r.get("DOO")


def test_zvector_solves_the_response_equations_at_20o200v():
    """Size-independent check at 20o/200v (4000 VO pairs): the Z-vector the
    GPU converged satisfies RPAA's fixed point (eps_a - eps_i) X - A X +
    B conj X = XVO, with A(ai,ck) = <ak||ci>, B(ai,ck) = <ac||ik> unpacked
    from the handle's own VOVO / VVOO classes (torch on the GPU), to the
    convergence threshold; Y = conj X."""
    no, nv = 20, 200
    s = synth.make_system(no, nv, seed=5, scale=0.01, device="cuda")
    r = solver_for(s)
    nit, rms = r.mp2_density(maxit=100, maxdim=8, ntol=11)
    assert rms < 1e-11
    dev = torch.device("cuda")
    nvo = no * nv
    vovo = torch.from_numpy(r.get("VOVO")).to(dev).reshape(no, nv, no, nv)   # [i][c][k][a]
    A = vovo.permute(0, 3, 2, 1).reshape(nvo, nvo)                          # rows (i,a), cols (k,c)
    del vovo
    p, q = synth.pair_index(nv)
    u, w = synth.pair_index(no)
    vv = torch.from_numpy(r.get("VVOO")).to(dev).reshape(len(u), len(p)).T   # [(a>c), (i>k)]
    D = torch.zeros(nv, nv, no, no, dtype=torch.complex128, device=dev)
    P, Q = torch.from_numpy(p).to(dev)[:, None], torch.from_numpy(q).to(dev)[:, None]
    U, W = torch.from_numpy(u).to(dev)[None, :], torch.from_numpy(w).to(dev)[None, :]
    D[P, Q, U, W] = vv
    D[Q, P, U, W] = -vv
    D[P, Q, W, U] = -vv
    D[Q, P, W, U] = vv
    B = D.permute(2, 0, 3, 1).reshape(nvo, nvo)                              # <ac||ik>
    del D, vv
    eps = torch.from_numpy(r.get_eps()).to(dev)
    de = (eps[no:][None, :] - eps[:no][:, None]).reshape(nvo).to(torch.complex128)
    X = torch.from_numpy(r.get("DVO")).to(dev)
    rhs = torch.from_numpy(r.get("XVO")).to(dev)
    res = de * X - A @ X + B @ torch.conj(X) - rhs
    assert float(res.abs().max()) < 1e-9 * float(rhs.abs().max()), float(res.abs().max())
    assert np.array_equal(r.get("DOV"), np.conj(r.get("DVO")))
