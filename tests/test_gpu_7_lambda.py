"""GPU checks of the CCSD Lambda equations and the response density.

* The adjoint sweep (rcc_lambda_vjp) is the exact transpose of the
  derivative of the iteration's own right-hand side: for random Lam and U,
  Lam . (dS/dT U) = (Lam . dS/dT) . U, with dS/dT U taken from the forward
  term sequence at six points T + w^k U on the unit circle (S is a
  polynomial of degree 4 in T, so that first Fourier coefficient is exact).
* Lam agrees with the oracle's dense solve (oracle/lambda_oracle.py).
* The density's tr(D V) agrees with the relccsd reference's finite-field
  gradient (tests/golden/lambda_*.npz, fourth-order central differences
  of converged reference energies) and with the GPU's own central
  differences, also through the torch.autograd.Function.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from molfdir_amd import synth  # noqa: E402
from molfdir_amd.autograd import PerturbedCCSD, ccsd_energy  # noqa: E402
from molfdir_amd.relccsd import RelCCSD  # noqa: E402
from oracle import lambda_oracle as lo  # noqa: E402

from _golden import load, names, system_for  # noqa: E402

TERMS = ("hointm", "hvintm", "hovintm", "t1eqns", "gointm", "gvintm", "aintm", "hintm",
         "t2eqns", "bintm")


def rhs_at(r, t1, t2):
    r.set_amplitudes(t1, t2)
    for t in TERMS:
        r.term(t)
    return np.concatenate([r.get("S1"), r.get("S2")])


def jvp(r, t1, t2, u1, u2, npts=6):
    w = np.exp(2j * np.pi * np.arange(npts) / npts)
    acc = 0
    for k in range(npts):
        acc = acc + np.conj(w[k]) * rhs_at(r, t1 + w[k] * u1, t2 + w[k] * u2)
    return acc / npts


def rand(rng, n, s):
    return s * (rng.normal(size=n) + 1j * rng.normal(size=n))


SYSTEMS = [
    # no, nv, seed, scale, fock_perturb, frozen
    (4, 10, 1, 0.05, 0.0, None),
    (5, 14, 4, 0.05, 0.05, None),
    (6, 40, 1, 0.05, 0.02, None),
    (3, 9, 7, 0.05, 0.03, [1, 0, 0]),
    (2, 5, 6, 0.05, 0.05, None),
]


@pytest.mark.parametrize("no,nv,seed,scale,fp,frozen", SYSTEMS)
def test_adjoint_sweep_is_the_transpose_of_the_forward_derivative(no, nv, seed, scale, fp,
                                                                  frozen):
    s = synth.make_system(no, nv, seed=seed, scale=scale, fock_perturb=fp)
    r = RelCCSD(no, nv, frozen=frozen)
    r.set_system(s)
    rng = np.random.default_rng(seed)
    t1, t2 = rand(rng, r.nt1, 0.05), rand(rng, r.nt2, 0.05)
    u1, u2 = rand(rng, r.nt1, 0.05), rand(rng, r.nt2, 0.05)
    l1, l2 = rand(rng, r.nt1, 1.0), rand(rng, r.nt2, 1.0)
    ju = jvp(r, t1, t2, u1, u2)
    r.set_amplitudes(t1, t2)
    g1, g2 = r.lambda_vjp(l1, l2, with_energy=False)
    lhs = np.concatenate([l1, l2]) @ ju
    rhs = np.concatenate([g1, g2]) @ np.concatenate([u1, u2])
    scale_ = np.abs(np.concatenate([l1, l2])).sum() * np.abs(ju).max()
    assert abs(lhs - rhs) < 1e-13 * scale_, (lhs, rhs)
    # per block: T1 and T2 parts separately (catches compensating errors)
    ju1 = jvp(r, t1, t2, u1, 0 * u2)
    r.set_amplitudes(t1, t2)
    assert abs(np.concatenate([l1, l2]) @ ju1 - g1 @ u1) < 1e-13 * scale_
    # the energy part: dE/dT . U against CCENG's complex sum, differentiated
    # exactly (degree 2) on three circle points
    g1e, g2e = r.lambda_vjp(0 * l1, 0 * l2, with_energy=True)
    w = np.exp(2j * np.pi * np.arange(3) / 3)
    ref = sum(np.conj(w[k]) * eccsd_complex(s, t1 + w[k] * u1, t2 + w[k] * u2)
              for k in range(3)) / 3
    assert abs(g1e @ u1 + g2e @ u2 - ref) < 1e-13 * max(1.0, abs(ref))


def eccsd_complex(s, t1, t2):
    """zdotc(FVO, T1) + zdotc(VVOO, tau) (CCENG before the real part)."""
    from oracle import ccsd_oracle as co
    L = co.Layout(int(s["no"]), int(s["nv"]))
    tau = co.gettau(L.t1(t1), L.t2(t2))
    return np.vdot(s["FVO"].numpy(), t1) + np.vdot(s["VVOO"].numpy(), L.pt2(tau))


def converged(g, ntol=13):
    s = system_for(g)
    no, nv = int(g["params"][0]), int(g["params"][1])
    r = RelCCSD(no, nv, maxit=100, maxdim=8, ntol=ntol)
    r.set_system(s)
    out = r.cceqn()
    assert out["rows"][-1][2] < 10.0 ** -ntol
    return r, s, out


SMALL = ["lam_3o8v", "lam_4o12v", "lam_5o16v"]


@pytest.mark.parametrize("name", SMALL)
def test_lambda_matches_dense_oracle(name):
    g = load("lambda", name)
    r, s, out = converged(g)
    assert abs(out["eccsd"] - float(g["e0"])) < 1e-10
    nit, rms = r.lambda_solve()
    assert rms < 1e-12, (nit, rms)
    l1, l2 = r.get_lambda()
    t1, t2 = r.get_amplitudes()
    snp = {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in s.items()}
    o1, o2 = lo.solve_lambda(snp, t1, t2)
    assert np.abs(l1 - o1).max() < 1e-9
    assert np.abs(l2 - o2).max() < 1e-9


def test_density_matches_dense_oracle():
    g = load("lambda", "lam_3o8v")
    r, s, _ = converged(g)
    r.lambda_solve()
    D = r.density()
    t1, t2 = r.get_amplitudes()
    l1, l2 = r.get_lambda()
    snp = {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in s.items()}
    ref = lo.density(snp, t1, t2, l1, l2)
    assert np.abs(D - D.conj().T).max() < 1e-15
    assert np.abs(D - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("name", names("lambda"))
def test_gradient_matches_reference_finite_field(name):
    g = load("lambda", name)
    r, s, _ = converged(g, ntol=12)
    nit, rms = r.lambda_solve()
    assert rms < 1e-12
    D = r.density()
    grad = float(np.real(np.trace(D @ g["V"])))
    fd = float(g["fd4"])
    assert abs(grad - fd) <= 2e-6 * abs(fd) + 2e-11, (grad, fd)


def test_autograd_function_matches_gpu_central_differences():
    g = load("lambda", "lam_6o40v")
    s = system_for(g)
    prob = PerturbedCCSD(s, maxit=100, ntol=12)
    n = g["V"].shape[0]
    # at V = 0: the reference's energy, and the gradient along the fixture's V
    V0 = torch.zeros((n, n), dtype=torch.complex128, requires_grad=True)
    e = ccsd_energy(V0, prob)
    e.backward()
    assert abs(e.item() - float(g["e0"])) < 1e-10
    d = float(np.real(np.sum(np.conj(V0.grad.numpy()) * g["V"])))
    assert abs(d - float(g["fd4"])) <= 2e-6 * abs(float(g["fd4"])) + 2e-11
    # at a perturbed point V1 (not Hermitian: E uses its Hermitian part), along W
    rng = np.random.default_rng(4)
    V1n = g["V"] + 0.005 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    W = lo.hermitian_direction(n, 99)
    V1 = torch.from_numpy(V1n).requires_grad_(True)
    e1 = ccsd_energy(V1, prob)
    e1.backward()
    grad = V1.grad.numpy()
    assert np.abs(grad - grad.conj().T).max() < 1e-15
    d = float(np.real(np.sum(np.conj(grad) * W)))
    h = 1e-4
    fd = (prob.energy(V1n + h * W) - prob.energy(V1n - h * W)) / (2 * h)
    assert abs(d - fd) <= 2e-6 * abs(fd) + 2e-11, (d, fd)


def test_frozen_core_gradient_matches_gpu_central_differences():
    s = synth.make_system(4, 12, seed=5, scale=0.03, fock_perturb=0.03)
    frozen = [1, 0, 0, 0]
    prob = PerturbedCCSD(s, maxit=100, ntol=12, frozen=frozen)
    V = lo.hermitian_direction(16, 31)
    e, D = prob.energy_and_density(np.zeros((16, 16)))
    d = float(np.real(np.trace(D @ V)))
    h = 1e-4
    fd = (prob.energy(h * V) - prob.energy(-h * V)) / (2 * h)
    assert abs(d - fd) <= 2e-6 * abs(fd) + 2e-11, (d, fd)


@pytest.mark.parametrize("name", ["s4_5o12v", "s8_8o24v"])
def test_symmetric_handle_lambda_and_density_equal_c1_handle(name):
    # the symmetry-blocked handle (grouped transposed ladder over the irrep
    # blocks) and a C1 handle on the same integrals give the same Lambda and
    # density
    from _golden import layout_for
    g = load("symccsd", name)
    s = system_for(g)
    lay = layout_for(g)
    no, nv = int(g["params"][0]), int(g["params"][1])
    c1 = RelCCSD(no, nv, maxit=100, maxdim=8, ntol=12)
    c1.set_system(s)
    sy = RelCCSD(no, nv, maxit=100, maxdim=8, ntol=12,
                 sym=(g["sym_no"].tolist(), g["sym_nv"].tolist()))
    sy.set_fock(lay.to_sym("FOO", s["FOO"].numpy()), lay.to_sym("FVO", s["FVO"].numpy()),
                lay.to_sym("FVV", s["FVV"].numpy()), s["eps"])
    for c in ("OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV"):
        sy.set_ints(c, lay.to_sym(c, s[c].numpy()))
    for r in (c1, sy):
        out = r.cceqn()
        assert out["rows"][-1][2] < 1e-12
        nit, rms = r.lambda_solve()
        assert rms < 1e-12, (nit, rms)
    a1, a2 = c1.get_lambda()
    b1, b2 = sy.get_lambda()
    assert np.abs(lay.to_c1("T1", b1) - a1).max() < 1e-9
    assert np.abs(lay.to_c1("T2", b2) - a2).max() < 1e-9
    assert np.abs(sy.density() - c1.density()).max() < 1e-10
