"""The benchmark workloads at their full sizes, as GPU tests.

* 30o/300v C1 (factor std 0.005, at which the iteration converges): CCSD
  converged to rms < 1e-9 in FP64 and in each split-precision mode (i8x4,
  i8x5, bf16); every split mode's converged Ec within 1 uEh of FP64's.
* 20o/200v C1 (std 0.005): converged to rms < 1e-8 with the Gauss 3M
  complex product and with the classical 4M one; Ec to 1e-10 Eh, T to 1e-9.
* The 64-system symmetric mix (o ~ U{4..16}, v ~ U{40..160}, 8 irreps,
  Dirichlet(2) proportions): four iterations of every system
  - through the stream batch, eager and replayed from per-system graphs:
    bit for bit the per-system eager iterations (same kernels, same K
    splits, same reductions);
  - through the compiled grouped batch, graph-replayed: bit for bit its own
    eager replay.  Against the per-system iterations the grouped batch
    agrees to rounding only: over 64 systems the compiler merges each long
    product's K split into fewer, longer chunks (one launch of about 16K
    workgroups instead of 64 chip-filling ones), which changes the order of
    the k sums.
* 20o/200v C1 (std 0.01): the Lambda response density's tr(D V) along a
  seeded Hermitian V against the central difference (h = 1e-4) of converged
  CCSD energies.
"""
import gc

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from molfdir_amd import synth  # noqa: E402
from molfdir_amd.relccsd import RelCCSD  # noqa: E402

CLASSES = ("OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV")


def c1_handle(no, nv, seed, scale, **kw):
    s = synth.make_system(no, nv, seed=seed, scale=scale, device="cuda")
    r = RelCCSD(no, nv, **kw)
    r.set_fock(s["FOO"], s["FVO"], s["FVV"], s["eps"])
    for c in CLASSES:
        r.set_ints(c, s[c])
        s[c] = None
    del s
    torch.cuda.empty_cache()
    return r


def release(*handles):
    for r in handles:
        r.close()
    gc.collect()
    torch.cuda.empty_cache()


# ---------------------------------------------------------------- 30o/300v
CFG3 = dict(no=30, nv=300, seed=3, scale=0.005, ntol=9, maxit=40)


@pytest.fixture(scope="module")
def cfg3():
    print(f"cfg3: {CFG3}", flush=True)
    r = c1_handle(CFG3["no"], CFG3["nv"], CFG3["seed"], CFG3["scale"], maxdim=6,
                  maxit=CFG3["maxit"], ntol=CFG3["ntol"])
    base = r.cceqn()
    print(f"cfg3 fp64: Ec {base['eccsd']:.12f} nit {base['nit']} rms {base['rows'][-1][2]:.2e}",
          flush=True)
    yield r, base
    release(r)


def test_cfg3_fp64_converges(cfg3):
    r, base = cfg3
    assert base["nit"] < CFG3["maxit"]
    assert base["rows"][-1][2] < 10.0 ** -CFG3["ntol"]
    assert np.isfinite(base["eccsd"]) and base["eccsd"] < 0


@pytest.mark.parametrize("mode", ["i8x4", "i8x5", "bf16"])
def test_cfg3_split_mode_within_1uEh_of_fp64(cfg3, mode):
    r, base = cfg3
    if mode == "bf16":
        r.set_options(6, CFG3["maxit"], CFG3["ntol"], 1)
    else:
        r.set_options(6, CFG3["maxit"], CFG3["ntol"], 2)
        r.set_split_digits(int(mode[-1]))
    try:
        out = r.cceqn()
    finally:
        r.set_options(6, CFG3["maxit"], CFG3["ntol"], 0)
    d = abs(out["eccsd"] - base["eccsd"])
    print(f"cfg3 {mode}: Ec {out['eccsd']:.12f} nit {out['nit']} |dEc| {d:.2e}", flush=True)
    assert out["rows"][-1][2] < 10.0 ** -CFG3["ntol"]
    assert d < 1e-6


# ---------------------------------------------------------------- 20o/200v
def test_cfg2_3m_and_4m_products_converge_to_the_same_state():
    no, nv, seed, scale, ntol = 20, 200, 2, 0.005, 8
    print(f"cfg2: {no}o/{nv}v scale {scale} ntol {ntol}", flush=True)
    r = c1_handle(no, nv, seed, scale, maxdim=6, maxit=40, ntol=ntol)
    try:
        out, amps = {}, {}
        for three in (True, False):
            r.set_complex_product(three)
            out[three] = r.cceqn()
            amps[three] = r.get_amplitudes()
            print(f"cfg2 {'3M' if three else '4M'}: Ec {out[three]['eccsd']:.12f} "
                  f"nit {out[three]['nit']}", flush=True)
    finally:
        release(r)
    for three in (True, False):
        assert out[three]["rows"][-1][2] < 10.0 ** -ntol
    assert abs(out[True]["eccsd"] - out[False]["eccsd"]) < 1e-10
    assert np.abs(amps[True][0] - amps[False][0]).max() < 1e-9
    assert np.abs(amps[True][1] - amps[False][1]).max() < 1e-9


# ---------------------------------------------------------------- 64 systems
NIT_BATCH = 4


def batch_handles():
    sizes = synth.batch_sizes(64, 8, 4)
    hs = []
    for o, v, no_l, nv_l, sd in sizes:
        sym = (no_l, nv_l)
        d = synth.make_system(o, v, seed=sd, scale=0.005, device="cuda", sym=sym)
        maps = synth.block_maps(sym)
        r = RelCCSD(o, v, sym=sym, maxdim=6, maxit=1000, ntol=12)
        r.set_fock(*(synth.to_blocks(n, d[n], sym, maps).cpu() for n in ("FOO", "FVO", "FVV")),
                   d["eps"])
        for name in CLASSES:
            r.set_ints(name, synth.to_blocks(name, d[name], sym, maps).contiguous())
        del d
        hs.append(r)
    torch.cuda.synchronize()
    return sizes, hs


def run(hs, step):
    """MP2 guess on every handle, then NIT_BATCH iterations through `step`;
    returns the per-handle (nit, Ec, rms) rows and final amplitudes."""
    for r in hs:
        r.mp2()
    rows = [[] for _ in hs]
    for _ in range(NIT_BATCH):
        for i, row in enumerate(step(hs)):
            rows[i].append(row)
    return rows, [r.get_amplitudes() for r in hs]


def same_bits(a, b, i):
    """Handle i's rows and amplitudes in runs a and b, bit for bit."""
    return a[0][i] == b[0][i] and all(np.array_equal(x.view(np.uint64), y.view(np.uint64))
                                      for x, y in zip(a[1][i], b[1][i]))


def test_cfg4_64_system_batch_paths_against_per_system_iterations():
    sizes, hs = batch_handles()
    print(f"cfg4: {len(hs)} systems, o {min(s[0] for s in sizes)}-{max(s[0] for s in sizes)}, "
          f"v {min(s[1] for s in sizes)}-{max(s[1] for s in sizes)}", flush=True)
    try:
        eager = run(hs, lambda hs: [r.iterate() for r in hs])
        stream = run(hs, RelCCSD.iterate_batch)
        for r in hs:
            r.set_graph(True)
        # iteration 1 eager, 2 captured, 3 and 4 replayed
        stream_graph = run(hs, RelCCSD.iterate_batch)
        for r in hs:
            r.set_graph(False)
        grouped = run(hs, lambda hs: RelCCSD.iterate_batch(hs, grouped=True))
        plan = RelCCSD.batch_plan_info(hs)
        for r in hs:
            r.set_graph(True)
        # the compiled plan replayed, then captured, then replayed from the graph
        grouped_graph = run(hs, lambda hs: RelCCSD.iterate_batch(hs, grouped=True))
        print(f"cfg4 plan: {plan}", flush=True)
    finally:
        release(*hs)
    assert plan["compiled"]
    for i in range(len(hs)):
        assert same_bits(eager, stream, i), i
        assert same_bits(eager, stream_graph, i), i
        assert same_bits(grouped, grouped_graph, i), i
        for (n0, e0, r0), (n1, e1, r1) in zip(eager[0][i], grouped[0][i]):
            assert n0 == n1
            assert abs(e0 - e1) < 1e-12, (i, n0, e0, e1)
            # near convergence the rms is a difference of nearly equal
            # amplitudes: rounding of order 1e-18 Eh, hence the absolute floor
            assert abs(r0 - r1) < 1e-9 * r0 + 1e-16, (i, n0, r0, r1)
        for x, y in zip(eager[1][i], grouped[1][i]):
            assert np.abs(x - y).max() < 1e-12, i


# ---------------------------------------------------------------- Lambda gradient
def test_cfg5_lambda_gradient_against_central_differences():
    from molfdir_amd.autograd import PerturbedCCSD
    no, nv, seed, scale = 20, 200, 5, 0.01
    n = no + nv
    print(f"cfg5: {no}o/{nv}v scale {scale}", flush=True)
    sysd = synth.make_system(no, nv, seed=seed, scale=scale, device="cuda")
    prob = PerturbedCCSD(sysd, maxit=120, maxdim=8, ntol=12)
    del sysd
    torch.cuda.empty_cache()
    try:
        prob.energy(np.zeros((n, n)))
        nit, rms = prob.r.lambda_solve()
        assert rms < 1e-12, (nit, rms)
        D = prob.r.density()
        V = synth.hermitian_perturbation(n, seed + 100, 0.01)
        g = float(np.real(np.trace(D @ V)))
        h = 1e-4
        fd = (prob.energy(h * V) - prob.energy(-h * V)) / (2 * h)
    finally:
        release(prob.r)
    print(f"cfg5: tr(DV) {g:.12e} central difference {fd:.12e} rel {abs(g - fd) / abs(fd):.2e}",
          flush=True)
    # the difference of two energies converged to rms 1e-12 carries ~1e-14 Eh
    # of noise per energy: ~1e-10 on the quotient, ~3e-7 relative here
    assert abs(g - fd) < 5e-6 * abs(fd)
