"""The plain FP64 GPU path against relccsd's CPU CCSD at C1 20o/150v, the
largest 20-occupied size relccsd's file layer can hold (ccunix.c wrabsf /
rdabsf compute byte offsets in 32-bit ints: 2 GiB per class file; VVVV at
20o/200v would be 6.3 GB).  The same seeded system (factor std 0.005, at
which the iteration converges), converged to NTOL 8 with DIIS dimension 6.
The reference run is held in tests/golden/cfg2ref_cfg2_20o150v.npz
(energies, the per-iteration table, T1, the 4096 largest T2 elements with
their indices, the T2 norm and 16 seeded +-1 projections of T2, since the
whole T2 is 2.1 M elements).

* MP2 and CCSD energies to 1e-10 Eh, the same iteration count, every
  iteration's energy to 1e-10 and RMS to 1e-9 relative;
* T1 and the largest T2 elements to 1e-9 max-abs, the T2 norm to 1e-12
  relative, the projections to 1e-9 per unit weight.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from molfdir_amd import synth  # noqa: E402
from molfdir_amd.relccsd import RelCCSD  # noqa: E402

from _golden import GOLDEN  # noqa: E402

REF = os.path.join(GOLDEN, "cfg2ref_cfg2_20o150v.npz")


@pytest.mark.skipif(not os.path.exists(REF), reason="20o/150v reference run not present")
def test_fp64_path_matches_relccsd_cpu_at_20o150v():
    g = np.load(REF, allow_pickle=False)
    no, nv, seed, maxit, maxdim, ntol = (int(x) for x in g["params"])
    print(f"cfg2: {no}o/{nv}v scale {float(g['scale'])} ntol {ntol}", flush=True)
    s = synth.make_system(no, nv, seed=seed, scale=float(g["scale"]), device="cuda")
    r = RelCCSD(no, nv, maxdim=maxdim, maxit=maxit, ntol=ntol)
    r.set_system(s)
    del s
    torch.cuda.empty_cache()
    o = r.cceqn()
    t1, t2 = r.get_amplitudes()
    r.close()
    assert abs(o["emp2"] - float(g["emp2"])) < 1e-10
    assert abs(o["eccsd"] - float(g["eccsd"])) < 1e-10
    ref_rows = g["iterations"][1:]          # the reference table starts at iteration 0
    assert o["nit"] == int(g["nit"]) == len(ref_rows)
    for (nit, e, rms), ref in zip(o["rows"], ref_rows):
        assert abs(e - ref[1]) < 1e-10, (nit, e, ref[1])
        assert abs(rms - ref[2]) < 1e-9 * max(1.0, ref[2]), (nit, rms, ref[2])
    assert np.abs(t1 - g["T1"]).max() < 1e-9
    top = g["T2_top_index"]
    assert np.abs(t2[top] - g["T2_top"]).max() < 1e-9
    assert abs(np.linalg.norm(t2) - float(g["T2_norm"])) < 1e-12 * float(g["T2_norm"])
    rng = np.random.Generator(np.random.PCG64(2024))
    w = rng.integers(0, 2, (g["T2_proj"].size, t2.size), dtype=np.int8) * 2 - 1
    proj = np.array([np.dot(wk.astype(np.float64), t2) for wk in w])
    assert np.abs(proj - g["T2_proj"]).max() < 1e-9 * np.sqrt(t2.size)
