"""Loading of the committed golden fixtures (outputs of the compiled relccsd
reference, see tests/golden/make_golden.py)."""
import glob
import os

import numpy as np

from molfdir_amd import synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names(kind):
    return sorted(os.path.basename(p)[len(kind) + 1:-4]
                  for p in glob.glob(os.path.join(GOLDEN, f"{kind}_*.npz")))


def load(kind, name):
    z = np.load(os.path.join(GOLDEN, f"{kind}_{name}.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def system_for(g, device="cpu"):
    p = g["params"]
    no, nv, seed = int(p[0]), int(p[1]), int(p[2])
    sym = None
    if "sym_no" in g:
        sym = (g["sym_no"].tolist(), g["sym_nv"].tolist())
    s = synth.make_system(no, nv, seed=seed, scale=float(g["scale"]),
                          fock_perturb=float(g["fock_perturb"]), device=device, sym=sym)
    if "nfroz" in g:
        s["nfroz"] = g["nfroz"].tolist()
    return s


def frozen_mask_for(g):
    """Frozen mask over the occupied spinors (C1 working order, irrep-major)
    of a fixture with NFROZ, else None."""
    if "nfroz" not in g:
        return None
    no_l = g["sym_no"].tolist() if "sym_no" in g else [int(g["params"][0])]
    mask = []
    for n, f in zip(no_l, g["nfroz"].tolist()):
        mask += [1] * f + [0] * (n - f)
    return mask


def fixture_amplitudes(g):
    """The random T1/T2 (C1 layout) a terms / symterms / frzterms fixture was
    evaluated on."""
    no, nv, tseed = int(g["params"][0]), int(g["params"][1]), int(g["params"][3])
    if "sym_no" in g:
        return symmetric_amplitudes(layout_for(g), tseed)
    return random_amplitudes(no, nv, tseed)


def layout_for(g):
    """oracle.symtab layout of a symmetric fixture."""
    from oracle.symtab import SymLayout
    return SymLayout(g["sym_no"].tolist(), g["sym_nv"].tolist())


# layout of each per-term output
TERM_LAYOUT = {"HO": "FOO", "GO": "FOO", "HV": "FVV", "GV": "FVV", "HOV": "T1", "S1": "T1",
               "S1_DENOM": "T1", "AZ": "OOOO", "H": "VOVO", "S2_T2EQNS": "T2",
               "S2_BINTM": "T2", "S2_DENOM": "T2"}


def symmetric_amplitudes(lay, seed):
    """Random T1/T2 with the layout's symmetry, in C1 layout (as the
    symmetric fixtures were made)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n1, n2 = lay.size("T1"), lay.size("T2")
    t1 = 0.05 * (rng.normal(size=n1) + 1j * rng.normal(size=n1))
    t2 = 0.05 * (rng.normal(size=n2) + 1j * rng.normal(size=n2))
    return lay.to_c1("T1", t1), lay.to_c1("T2", t2)


def fingerprint(sysd):
    keys = ("eps", "FOO", "FVO", "FVV", "OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV")
    out = []
    for k in keys:
        v = np.asarray(sysd[k].cpu().numpy())
        out.append([np.abs(v).sum(), np.real(v).sum(), np.imag(v).sum() if np.iscomplexobj(v) else 0.0])
    return np.asarray(out)


def random_amplitudes(no, nv, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    noot, nvvt = no * (no - 1) // 2, nv * (nv - 1) // 2
    t1 = 0.05 * (rng.normal(size=no * nv) + 1j * rng.normal(size=no * nv))
    t2 = 0.05 * (rng.normal(size=noot * nvvt) + 1j * rng.normal(size=noot * nvvt))
    return t1, t2
