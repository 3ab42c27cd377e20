"""rcc_zgemm_device_noalloc, the allocation-free form of rcc_zgemm_device:
called through the C ABI (ctypes), its products equal those of
rcc_zgemm_device on the same operands -- bit for bit when both take the same
K split (the autotuner off: the same kernels sum in the same order), and to
the FP64 GEMM bar (1e-13 relative) in any case -- on two calls in a row, the
second one smaller, so the scratch still holds the first call's partial sums
when the second one runs.  A split whose partial sums exceed the 32 MiB
scratch is refused with an error."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from molfdir_amd.relccsd import RelCCSD, RccError  # noqa: E402

OP_N, OP_T, OP_C = 0, 1, 2


def operand(rows, cols, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = torch.randn(rows, cols, dtype=torch.float64, generator=g) + \
        1j * torch.randn(rows, cols, dtype=torch.float64, generator=g)
    # column-major storage: a (cols, rows) row-major tensor
    return x.T.contiguous().to(torch.complex128).cuda()


def one_product(r, opa, opb, m, n, k, seed, ksplit=0):
    ar, ac = (m, k) if opa == OP_N else (k, m)
    br, bc = (k, n) if opb == OP_N else (n, k)
    A, B = operand(ar, ac, seed), operand(br, bc, seed + 1)
    C0 = operand(m, n, seed + 2)
    alpha, beta = complex(0.7, -0.3), complex(0.25, 0.5)
    want, got = C0.clone(), C0.clone()
    r.zgemm_device(opa, opb, m, n, k, alpha, A, ar, B, br, beta, want, m)
    r.zgemm_device_noalloc(opa, opb, m, n, k, alpha, A, ar, B, br, beta, got, m, ksplit)
    r.synchronize()
    return want.cpu().numpy(), got.cpu().numpy()


@pytest.mark.parametrize("three_mult", [True, False])
def test_noalloc_product_equals_the_allocating_entry_point(three_mult):
    r = RelCCSD(2, 4)
    r.set_complex_product(three_mult)
    r.set_autotune(False)
    # first call: 300 x 200 output (15 tiles) with a long k, split over k
    want, got = one_product(r, OP_N, OP_C, 300, 200, 4000, seed=11)
    assert np.array_equal(want, got)
    # second call, smaller, other ops: the scratch holds the first call's sums
    want, got = one_product(r, OP_T, OP_N, 70, 50, 1500, seed=21)
    assert np.array_equal(want, got)
    # an explicit split sums in another order: the FP64 bar
    want, got = one_product(r, OP_N, OP_N, 64, 96, 2000, seed=31, ksplit=7)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


def test_noalloc_refuses_a_split_beyond_the_scratch():
    r = RelCCSD(2, 4)
    m = n = 1024                         # 16 MiB of partial sums per split
    A = torch.zeros(m * 64, dtype=torch.complex128, device="cuda")
    C = torch.zeros(m * n, dtype=torch.complex128, device="cuda")
    with pytest.raises(RccError, match="scratch"):
        r.zgemm_device_noalloc(OP_N, OP_N, m, n, 64, complex(1, 0), A, m, A, 64,
                               complex(0, 0), C, m, ksplit=4)
    # the same product without a split needs no scratch
    r.zgemm_device_noalloc(OP_N, OP_N, m, n, 64, complex(1, 0), A, m, A, 64,
                           complex(0, 0), C, m, ksplit=1)
    r.synchronize()


def test_noalloc_is_declared_in_the_public_header():
    from molfdir_amd import relccsd
    assert "rcc_zgemm_device_noalloc" in relccsd.declared_symbols()
    lib = ctypes.CDLL(relccsd.LIB_PATH)
    assert hasattr(lib, "rcc_zgemm_device_noalloc")
