"""GPU checks of the split complex GEMMs and of the CCSD iteration run with
them: split-BF16 (Re, Im and Re+Im as three BF16 slices each on
v_mfma_f32_16x16x32_bf16, FP32 accumulation over 128-term chunks; precision
1) and split-INT8 (4 or 5 signed 7-bit digits of Re, Im and Re+Im on
v_mfma_i32_16x16x64_i8, exact int32 accumulation; precision 2).

Exact checks: integer operands small enough that every slice product and
every 128-term FP32 sum is exact must give numpy's product bit for bit (pins
the A/B/D lane layouts, chunk loop and exponent handling); power-of-two
rescaling of the operands must rescale the result bit for bit (pins the
per-row exponents).  Rounding checks: random operands to an rms-relative
error of 2^-17 (FP32 accumulation sets the floor), which a missing
first-order slice product exceeds by orders of magnitude; k = 1 products of
values whose Re, Im and Re+Im are each built from exactly one, two or three
BF16 slices must come out exact, which pins every retained slice product.
The CCSD trajectories are compared with the FP64 reference fixtures.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from molfdir_amd.relccsd import RelCCSD  # noqa: E402

from _golden import load, names, system_for  # noqa: E402


def dev(x):   # column-major flat copy on the GPU
    return torch.from_numpy(np.asarray(x, np.complex128).reshape(-1, order="F").copy()).to("cuda")


def split_gemm(A, B, C=None, alpha=1.0, beta=0.0):
    m, k = A.shape
    n = B.shape[1]
    tC = dev(np.zeros((m, n)) if C is None else C)
    r = RelCCSD(2, 4)
    r.zgemm_split_device(m, n, k, complex(alpha), dev(A), m, dev(B), k, complex(beta), tC, m)
    torch.cuda.synchronize()
    return tC.cpu().numpy().reshape(n, m).T


SHAPES = [(16, 16, 32), (64, 64, 128), (37, 91, 53), (5, 3, 3000), (130, 70, 1), (257, 129, 300)]


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_split_gemm_exact_on_small_integers(m, n, k):
    rng = np.random.default_rng(m + 3 * n + 7 * k)
    A = rng.integers(-64, 65, (m, k)) + 1j * rng.integers(-64, 65, (m, k))
    B = rng.integers(-64, 65, (k, n)) + 1j * rng.integers(-64, 65, (k, n))
    got = split_gemm(A, B)
    assert np.array_equal(got, A @ B)


def _sliced(rng, shape, terms):
    """Values sign * 2^f * (1 + a 2^-9 + b 2^-18) with a, b in {-1, 0, 1}
    (terms 2: b = 0): exactly the sum of that many BF16 slices."""
    a = rng.integers(-1, 2, shape) * 2.0 ** -9
    b = rng.integers(-1, 2, shape) * 2.0 ** -18 if terms == 3 else 0.0
    f = 2.0 ** rng.integers(-3, 1, shape)
    sgn = rng.choice([-1.0, 1.0], shape)
    return sgn * f * (1 + a + b)


def _pow2(rng, shape):
    return rng.choice([-1.0, 1.0], shape) * 2.0 ** rng.integers(-3, 1, shape)


@pytest.mark.parametrize("case", ["A3xB1", "A1xB3", "A2xB2"])
def test_split_gemm_slice_products_exact(case):
    # k = 1: each output is one complex product whose slice products and FP32
    # sums are exact, so every retained slice product (leading, first and
    # second order) must be present exactly once
    rng = np.random.default_rng(len(case) + ord(case[1]) + 17 * ord(case[4]))
    m, n = 100, 37
    mk = {"A3xB1": (3, 1), "A1xB3": (1, 3), "A2xB2": (2, 2)}[case]

    def op(shape, terms):
        # Im = s Re with s in {-1, 0, 1}: Re, Im and Re + Im (0, x or 2x) are
        # all sums of exactly `terms` BF16 slices
        x = _pow2(rng, shape) if terms == 1 else _sliced(rng, shape, terms)
        return x * (1 + 1j * rng.integers(-1, 2, shape))

    A = op((m, 1), mk[0])
    B = op((1, n), mk[1])
    got = split_gemm(A, B)
    assert np.array_equal(got, A @ B)


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_split_gemm_random_accuracy(m, n, k):
    rng = np.random.default_rng(11 * m + n + k)
    A = rng.normal(size=(m, k)) + 1j * rng.normal(size=(m, k))
    B = rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))
    C = rng.normal(size=(m, n)) + 1j * rng.normal(size=(m, n))
    alpha, beta = 0.7 - 0.2j, -0.3 + 0.5j
    ref = alpha * (A @ B) + beta * C
    got = split_gemm(A, B, C, alpha, beta)
    scale = np.sqrt(k) * np.abs(alpha) * np.sqrt(np.mean(np.abs(A) ** 2) * np.mean(np.abs(B) ** 2))
    rel = np.abs(got - ref).max() / scale
    assert rel < 2.0 ** -17, rel


def test_split_gemm_power_of_two_scaling_is_exact():
    rng = np.random.default_rng(5)
    A = rng.normal(size=(100, 300)) + 1j * rng.normal(size=(100, 300))
    B = rng.normal(size=(300, 90)) + 1j * rng.normal(size=(300, 90))
    base = split_gemm(A, B)
    scaled = split_gemm(A * 2.0 ** 40, B * 2.0 ** -70)
    assert np.array_equal(scaled, base * 2.0 ** -30)


def test_split_gemm_block_scaling_keeps_small_blocks_accurate():
    # rows 64.. of A are 1e-12 smaller: their own scale block keeps them at
    # the same relative accuracy as the large rows
    rng = np.random.default_rng(9)
    A = rng.normal(size=(128, 256)) + 1j * rng.normal(size=(128, 256))
    A[64:] *= 1e-12
    B = rng.normal(size=(256, 64)) + 1j * rng.normal(size=(256, 64))
    got = split_gemm(A, B)
    ref = A @ B
    for rows in (slice(0, 64), slice(64, 128)):
        s = np.sqrt(256 * np.mean(np.abs(A[rows]) ** 2) * np.mean(np.abs(B) ** 2))
        assert np.abs(got[rows] - ref[rows]).max() / s < 2.0 ** -17


@pytest.mark.parametrize("tile,kparts,gm", [(128, 3, 1), (64, 8, 4), (64, 2, 2), (64, 1, 1),
                                             (128, 1, 8)])
def test_split_gemm_tiles_ksplit_and_band_exact(monkeypatch, tile, kparts, gm):
    # both tiles (128 x 128 with three LDS buffers, 128 x 64 with two), the
    # K-split (partial sums reduced in a fixed order) and the tile-row bands
    # keep exact integer products exact, alpha and beta included
    monkeypatch.setenv("RCC_BF_TILE", str(tile))
    monkeypatch.setenv("RCC_BF_KPARTS", str(kparts))
    monkeypatch.setenv("RCC_BF_GM", str(gm))
    rng = np.random.default_rng(kparts + 10 * gm + tile)
    m, n, k = 300, 270, 1100
    A = rng.integers(-64, 65, (m, k)) + 1j * rng.integers(-64, 65, (m, k))
    B = rng.integers(-64, 65, (k, n)) + 1j * rng.integers(-64, 65, (k, n))
    C = rng.integers(-64, 65, (m, n)) + 1j * rng.integers(-64, 65, (m, n))
    got = split_gemm(A, B, C, 2.0 - 1.0j, -1.0 + 0.5j)
    assert np.array_equal(got, (2.0 - 1.0j) * (A @ B) + (-1.0 + 0.5j) * C)


@pytest.mark.parametrize("name", [x for x in names("ccsd") if x != "cfg1_6o40v"])
def test_split_bf16_trajectory_at_fp32_level(name):
    # FP32 accumulation bounds the BF16 scheme at a relative 1e-6 of Ec.  The
    # 6o/40v fixture is left out: its 20 unconverged iterations amplify any
    # perturbation (1e-6 at iteration 6 becomes 5e-4 by iteration 8)
    g = load("ccsd", name)
    maxit, maxdim, ntol = (int(x) for x in g["params"][3:6])
    s = system_for(g)
    nit = int(g["nit"])
    r = RelCCSD(int(s["no"]), int(s["nv"]), maxdim=maxdim, maxit=nit, ntol=ntol, precision=1)
    r.set_system(s)
    emp2 = r.mp2()
    assert abs(emp2 - float(g["emp2"])) < 1e-10   # MP2 has no split products
    rows = [r.iterate() for _ in range(nit)]
    for (it, e, rms), ref in zip(rows, g["iterations"][1:]):
        assert abs(e - ref[1]) < 1e-6 * abs(ref[1]), (it, e, ref[1])
    t1, t2 = r.get_amplitudes()
    assert np.abs(t1 - g["T1"]).max() < 1e-6
    assert np.abs(t2 - g["T2"]).max() < 1e-6


# ---------------------------------------------------------------- split-INT8
def i8_gemm(digits, A, B, C=None, alpha=1.0, beta=0.0):
    m, k = A.shape
    n = B.shape[1]
    tC = dev(np.zeros((m, n)) if C is None else C)
    r = RelCCSD(2, 4)
    r.zgemm_i8_device(digits, m, n, k, complex(alpha), dev(A), m, dev(B), k, complex(beta), tC, m)
    torch.cuda.synchronize()
    return tC.cpu().numpy().reshape(n, m).T


I8_SHAPES = SHAPES + [(70, 33, 8300)]   # 8300: two scale chunks


def _ints(rng, shape, bound):
    return rng.integers(-bound, bound + 1, shape) + 1j * rng.integers(-bound, bound + 1, shape)


@pytest.mark.parametrize("digits", [4, 5])
@pytest.mark.parametrize("m,n,k", I8_SHAPES)
def test_i8_gemm_exact_on_small_integers(digits, m, n, k):
    # |Re|, |Im| <= 31 keeps Re, Im and Re+Im single-digit: every product exact
    rng = np.random.default_rng(m + 3 * n + 7 * k + digits)
    A, B = _ints(rng, (m, k), 31), _ints(rng, (k, n), 31)
    assert np.array_equal(i8_gemm(digits, A, B), A @ B)


@pytest.mark.parametrize("digits", [4, 5])
@pytest.mark.parametrize("case", ["A1xBS", "ASxB1", "A2xBS-1"])
def test_i8_gemm_digit_pairs_exact(digits, case):
    # integers that need exactly the stated number of digits on one side: every
    # retained digit pair (i, j), i + j < S, is then required and sufficient
    rng = np.random.default_rng(digits * 100 + ord(case[1]) + len(case))
    m, n, k = 48, 40, 200
    big = 2 ** (6 + 7 * (digits - 1) - 2)      # S digits: |Re|, |Im|, |Re+Im| < 2^(7S-1)
    mid = 2 ** (6 + 7 * (digits - 2) - 2)      # S-1 digits
    if case == "A1xBS":
        A, B = _ints(rng, (m, k), 31), _ints(rng, (k, n), big)
    elif case == "ASxB1":
        A, B = _ints(rng, (m, k), big), _ints(rng, (k, n), 31)
    else:
        A, B = _ints(rng, (m, k), 2 ** 11), _ints(rng, (k, n), mid)
    assert np.array_equal(i8_gemm(digits, A, B), A @ B)


@pytest.mark.parametrize("digits", [4, 5])
@pytest.mark.parametrize("m,n,k", I8_SHAPES)
def test_i8_gemm_random_accuracy(digits, m, n, k):
    rng = np.random.default_rng(11 * m + n + k + digits)
    A = rng.normal(size=(m, k)) + 1j * rng.normal(size=(m, k))
    B = rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))
    C = rng.normal(size=(m, n)) + 1j * rng.normal(size=(m, n))
    alpha, beta = 0.7 - 0.2j, -0.3 + 0.5j
    ref = alpha * (A @ B) + beta * C
    got = i8_gemm(digits, A, B, C, alpha, beta)
    scale = np.sqrt(k) * np.abs(alpha) * np.sqrt(np.mean(np.abs(A) ** 2) * np.mean(np.abs(B) ** 2))
    rel = np.abs(got - ref).max() / scale
    # representation and dropped digit pairs: about 2^-(7 S - 4) of the scale
    assert rel < 2.0 ** -(7 * digits - 7), rel


def test_i8_gemm_power_of_two_scaling_is_exact():
    rng = np.random.default_rng(5)
    A = rng.normal(size=(100, 300)) + 1j * rng.normal(size=(100, 300))
    B = rng.normal(size=(300, 90)) + 1j * rng.normal(size=(300, 90))
    base = i8_gemm(5, A, B)
    assert np.array_equal(i8_gemm(5, A * 2.0 ** 40, B * 2.0 ** -70), base * 2.0 ** -30)


def test_i8_gemm_row_scaling_keeps_small_rows_accurate():
    rng = np.random.default_rng(9)
    A = rng.normal(size=(128, 256)) + 1j * rng.normal(size=(128, 256))
    A[1::2] *= 1e-12                       # every other row: scale is per row
    B = rng.normal(size=(256, 64)) + 1j * rng.normal(size=(256, 64))
    got, ref = i8_gemm(5, A, B), A @ B
    for rows in (slice(0, None, 2), slice(1, None, 2)):
        s = np.sqrt(256 * np.mean(np.abs(A[rows]) ** 2) * np.mean(np.abs(B) ** 2))
        assert np.abs(got[rows] - ref[rows]).max() / s < 2.0 ** -28


@pytest.mark.parametrize("digits,tol", [(4, 1e-6), (5, 1e-8)])
@pytest.mark.parametrize("name", names("ccsd"))
def test_split_i8_trajectory_matches_fp64(name, digits, tol):
    g = load("ccsd", name)
    maxit, maxdim, ntol = (int(x) for x in g["params"][3:6])
    s = system_for(g)
    nit = int(g["nit"])
    r = RelCCSD(int(s["no"]), int(s["nv"]), maxdim=maxdim, maxit=nit, ntol=ntol, precision=2)
    r.set_split_digits(digits)
    r.set_system(s)
    assert abs(r.mp2() - float(g["emp2"])) < 1e-10
    if name == "cfg1_6o40v":
        # 20 unconverged iterations that amplify any perturbation (a 1e-10
        # difference grows past 1e-7 by iteration 8): compare the first 6
        nit = 6
    rows = [r.iterate() for _ in range(nit)]
    for (it, e, rms), ref in zip(rows, g["iterations"][1:]):
        assert abs(e - ref[1]) < tol * max(1.0, abs(ref[1])), (it, e, ref[1])
    if nit == int(g["nit"]):
        t1, t2 = r.get_amplitudes()
        assert np.abs(t1 - g["T1"]).max() < tol
        assert np.abs(t2 - g["T2"]).max() < tol


@pytest.mark.parametrize("m,n,k", [(200, 300, 700), (130, 129, 8300), (64, 64, 64), (300, 140, 1000)])
def test_i8_tile_shapes_give_identical_results(m, n, k, monkeypatch):
    # every tile shape (32 WM x 64 for WM = 2, 3, 4, and the 128 x 128 tile of
    # 8 waves, "8") sums the same products in the same order: bit-identical C
    rng = np.random.default_rng(m + n + k)
    A = rng.normal(size=(m, k)) + 1j * rng.normal(size=(m, k))
    B = rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))
    C = rng.normal(size=(m, n)) + 1j * rng.normal(size=(m, n))
    outs = []
    for wm in ("2", "3", "4", "8"):
        monkeypatch.setenv("RCC_I8_WM", wm)
        outs.append(i8_gemm(4, A, B, C, 0.7 - 0.2j, -0.3 + 0.5j))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
