"""The Lambda / response-density oracle (oracle/lambda_oracle.py) against the
relccsd reference's finite-field gradients (tests/golden/lambda_*.npz: the
reference's converged CCSD energy at f + lam V, lam = 0, +-h, +-2h) and
against identities the mathematics fixes."""
import numpy as np
import pytest

from oracle import ccsd_oracle as co
from oracle import lambda_oracle as lo

from _golden import load, names, system_for

# dense Jacobians: only the small fixtures
SMALL = [n for n in names("lambda") if n in ("lam_3o8v", "lam_4o12v", "lam_5o16v")]


def _np_system(g):
    s = system_for(g)
    return {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in s.items()}


@pytest.mark.parametrize("name", SMALL)
def test_oracle_gradient_matches_reference_finite_field(name):
    g = load("lambda", name)
    s = _np_system(g)
    # the reference's own converged amplitudes at lam = 0
    L1, L2 = lo.solve_lambda(s, g["T1"], g["T2"])
    d = lo.directional_derivative(s, g["T1"], g["T2"], L1, L2, g["V"])
    fd = float(g["fd4"])
    assert abs(d - fd) <= 2e-6 * abs(fd) + 2e-11, (d, fd)


def test_fixture_direction_is_the_seeded_one():
    for name in names("lambda"):
        g = load("lambda", name)
        n = int(g["params"][0] + g["params"][1])
        assert np.array_equal(g["V"], lo.hermitian_direction(n, int(g["params"][3])))


def test_first_coefficient_is_exact_for_the_residual_degree():
    # N circle points alias the order-1 coefficient with orders 1 + N, 1 + 2N ...
    # R has degree exactly 4 in T: N = 4, 6, 9 agree (no order 5), N = 3 picks
    # up the order-4 term and must differ
    g = load("lambda", "lam_3o8v")
    P = lo.Problem(_np_system(g))
    x = np.concatenate([g["T1"], g["T2"]])
    rng = np.random.default_rng(3)
    u = rng.normal(size=x.size) + 1j * rng.normal(size=x.size)
    c = {n: lo.first_coefficient(P.residual, x, u, n) for n in (3, 4, 6, 9)}
    scale = np.abs(c[6]).max()
    assert np.abs(c[4] - c[6]).max() < 1e-12 * scale
    assert np.abs(c[9] - c[6]).max() < 1e-12 * scale
    assert np.abs(c[3] - c[6]).max() > 1e-6 * scale


def test_density_is_hermitian_and_linear():
    g = load("lambda", "lam_3o8v")
    s = _np_system(g)
    L1, L2 = lo.solve_lambda(s, g["T1"], g["T2"])
    G = lo.density(s, g["T1"], g["T2"], L1, L2)
    assert np.abs(G - G.conj().T).max() == 0.0
    for seed in (1, 2):
        V = lo.hermitian_direction(G.shape[0], 100 + seed)
        d = lo.directional_derivative(s, g["T1"], g["T2"], L1, L2, V)
        assert abs(np.trace(G @ V).real - d) < 1e-12 * max(1.0, abs(d))
