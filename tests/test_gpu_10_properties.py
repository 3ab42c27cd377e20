"""One-electron properties: rcc_read_property (RDPROP ccints.F:569-764) and
rcc_expectation (EXPVAL ccgrad.F:261-347).

* The operator read from an MDPROP file equals the blocks relccsd's own
  RDPROP forms (AOO, AVV, AVO, AOV; tests/golden/prop_*.npz), bit for bit,
  and the reference-determinant expectation value equals relccsd's EXPVAL
  on zero densities (CCGRAD's SCF line).
* The unrelaxed CCSD expectation value minus the reference part equals the
  GPU's own central difference of the converged CCSD energy under
  f -> f + lam P (h = 1e-4, relative 2e-6), the same finite-field bar the
  Lambda density meets against the reference.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from molfdir_amd import mdcint, synth  # noqa: E402
from molfdir_amd.autograd import PerturbedCCSD  # noqa: E402
from molfdir_amd.relccsd import RelCCSD, RccError  # noqa: E402

from _golden import load, names  # noqa: E402


def files_for(g, d):
    no, nv, seed, pseed = (int(x) for x in g["params"])
    s = synth.make_system(no, nv, seed=seed, scale=float(g["scale"]),
                          fock_perturb=float(g["fock_perturb"]))
    P = synth.hermitian_perturbation(no + nv, pseed, 0.1)
    eps = s["eps"].numpy()
    mdcint.write_mrconee(str(d / "MRCONEE"), eps, np.ones(no + nv, np.int32), [no],
                         np.diag(eps).astype(np.complex128))
    mdcint.write_mdprop(str(d / "MDPROP"), {"OTHER": 2.0 * P, "PROPTEST": P})
    return s, P, no, nv


@pytest.mark.parametrize("name", names("prop"))
def test_rdprop_blocks_and_reference_expectation(name, tmp_path):
    g = load("prop", name)
    s, P, no, nv = files_for(g, tmp_path)
    r = RelCCSD(no, nv)
    M = r.read_property(tmp_path / "MRCONEE", tmp_path / "MDPROP", "PROPTEST", [no])
    assert np.array_equal(M, P)
    assert np.array_equal(M[:no, :no].T.reshape(-1), g["AOO"])
    assert np.array_equal(M[no:, no:].T.reshape(-1), g["AVV"])
    assert np.array_equal(M[no:, :no].T.reshape(-1), g["AVO"])
    assert np.array_equal(M[:no, no:].reshape(-1), g["AOV"])      # (A, I) of P(I, A)
    assert abs(r.expectation(M, level=0) - complex(g["expval"])) < 1e-15
    with pytest.raises(RccError):
        r.read_property(tmp_path / "MRCONEE", tmp_path / "MDPROP", "MISSING", [no])


def test_ccsd_expectation_is_the_finite_field_derivative(tmp_path):
    g = load("lambda", "lam_4o12v")
    no, nv = int(g["params"][0]), int(g["params"][1])
    from _golden import system_for
    s = system_for(g)
    P = synth.hermitian_perturbation(no + nv, 77, 0.1)
    prob = PerturbedCCSD(s, maxit=100, ntol=12)
    e0, _ = prob.energy_and_density(np.zeros((no + nv, no + nv)))
    val = prob.r.expectation(P, level=1)
    ref_part = np.trace(P[:no, :no])
    h = 1e-4
    fd = (prob.energy(h * P) - prob.energy(-h * P)) / (2 * h)
    corr = val - ref_part
    assert abs(corr.imag) < 1e-12
    assert abs(corr.real - fd) <= 2e-6 * abs(fd) + 2e-11, (corr, fd)
