"""GPU checks of the complex FP64-MFMA GEMM engine (through the C ABI)
against numpy, and of the v_mfma_f64_16x16x4_f64 lane layout it assumes."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from molfdir_amd.relccsd import RelCCSD  # noqa: E402


def test_mfma_f64_lane_layout():
    r = RelCCSD(2, 4)
    d = r.probe_mfma()
    a = np.array([[i + 16 * k for k in range(4)] for i in range(16)], dtype=np.float64)
    b = np.array([[(k + 1) * 1000 + j for j in range(16)] for k in range(4)], dtype=np.float64)
    ref = a @ b
    got = np.zeros((16, 16))
    for lane in range(64):
        for e in range(4):
            got[lane // 16 + 4 * e, lane % 16] = d[lane, e]
    assert np.array_equal(got, ref), d


@pytest.mark.parametrize("three_mult", [True, False])
@pytest.mark.parametrize("opa", [0, 1, 2])
@pytest.mark.parametrize("opb", [0, 1, 2])
@pytest.mark.parametrize("m,n,k", [(64, 64, 16), (37, 91, 53), (5, 3, 3000), (130, 70, 1),
                                   (257, 129, 77), (300, 30, 301), (129, 32, 40), (257, 33, 9),
                                   (1000, 1, 500)])
def test_zgemm_against_numpy(opa, opb, m, n, k, three_mult):
    rng = np.random.default_rng(m * 7 + n * 3 + k + 10 * opa + opb)

    def mat(r, c):
        return rng.normal(size=(r, c)) + 1j * rng.normal(size=(r, c))

    A = mat(m, k) if opa == 0 else mat(k, m)
    B = mat(k, n) if opb == 0 else mat(n, k)
    C = mat(m, n)
    alpha, beta = 0.7 - 0.2j, -0.3 + 0.5j

    def op(x, o):
        return x if o == 0 else (x.T if o == 1 else x.conj().T)

    ref = alpha * op(A, opa) @ op(B, opb) + beta * C
    def dev(x):   # column-major flat copy on the GPU
        return torch.from_numpy(x.reshape(-1, order="F").copy()).to("cuda")

    tA, tB, tC = dev(A), dev(B), dev(C)
    r = RelCCSD(2, 4)
    r.set_complex_product(three_mult)
    r.zgemm_device(opa, opb, m, n, k, alpha, tA, A.shape[0], tB, B.shape[0], beta, tC, m)
    got = tC.cpu().numpy().reshape(n, m).T
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()) * max(1, k) ** 0.5




@pytest.mark.parametrize("three_mult", [True, False])
@pytest.mark.parametrize("opa,opb", [(3, 0), (0, 3)])
@pytest.mark.parametrize("m,n,k", [(64, 64, 16), (37, 91, 53), (5, 3, 3000), (257, 129, 77)])
def test_zgemm_conjugate_operand_against_numpy(opa, opb, m, n, k, three_mult):
    # op 3: the operand conjugated, not transposed (the adjoint contractions)
    rng = np.random.default_rng(m * 5 + n + k + opa)

    def mat(r, c):
        return rng.normal(size=(r, c)) + 1j * rng.normal(size=(r, c))

    A, B, C = mat(m, k), mat(k, n), mat(m, n)
    alpha, beta = 0.4 + 0.9j, 0.25 - 0.5j
    ref = alpha * (A.conj() if opa == 3 else A) @ (B.conj() if opb == 3 else B) + beta * C

    def dev(x):
        return torch.from_numpy(x.reshape(-1, order="F").copy()).to("cuda")

    tA, tB, tC = dev(A), dev(B), dev(C)
    r = RelCCSD(2, 4)
    r.set_complex_product(three_mult)
    r.zgemm_device(opa, opb, m, n, k, alpha, tA, m, tB, k, beta, tC, m)
    got = tC.cpu().numpy().reshape(n, m).T
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()) * max(1, k) ** 0.5


@pytest.mark.parametrize("opa,opb,m,n,k,batch", [(2, 0, 7, 9, 40000, 1), (0, 0, 200, 20, 9000, 1),
                                                  (1, 0, 64, 20, 30000, 1)])
def test_autotuned_split_gemm_against_numpy(opa, opb, m, n, k, batch):
    # first call of the shape: the autotuner times its K-splits on a scratch
    # C before the real product, which must see the live C untouched
    rng = np.random.default_rng(m + n + k)

    def mat(r, c):
        return rng.normal(size=(r, c)) + 1j * rng.normal(size=(r, c))

    A = mat(m, k) if opa == 0 else mat(k, m)
    B = mat(k, n)
    C = mat(m, n)
    alpha, beta = 0.3 + 0.1j, 0.9 - 0.4j

    def op(x, o):
        return x if o == 0 else (x.T if o == 1 else x.conj().T)

    ref = alpha * op(A, opa) @ B + beta * C

    def dev(x):
        return torch.from_numpy(x.reshape(-1, order="F").copy()).to("cuda")

    r = RelCCSD(2, 4)
    r.set_autotune(True)
    before = r.autotune_entries()[0]
    for rep in range(2):            # tuned on the first call, cached on the second
        tA, tB, tC = dev(A), dev(B), dev(C)
        r.zgemm_device(opa, opb, m, n, k, alpha, tA, A.shape[0], tB, k, beta, tC, m)
        got = tC.cpu().numpy().reshape(n, m).T
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max() * k ** 0.5
    assert r.autotune_entries()[0] == before + 1
