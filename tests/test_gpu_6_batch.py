"""GPU checks of the graph-replayed iteration (rcc_set_graph) and of the
batched iteration of independent systems (rcc_iterate_batch).

The graph holds the term sequence HOINTM .. DENOM captured once; every
replay must reproduce the eager iteration bit for bit (the kernels are
deterministic), with the host's DIIS and energy work between replays.  The
batch runs several systems' term phases concurrently on their own streams
(or, grouped, with the four dominant products of all systems as one launch
per term); each system must follow its own reference trajectory.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from molfdir_amd.relccsd import RelCCSD  # noqa: E402

from _golden import load, names, system_for, layout_for  # noqa: E402

CLASSES = ("OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV")


def make_solver(kind, name, **kw):
    g = load(kind, name)
    s = system_for(g)
    maxit, maxdim, ntol = (int(x) for x in g["params"][3:6])
    if kind == "symccsd":
        lay = layout_for(g)
        r = RelCCSD(int(g["params"][0]), int(g["params"][1]), maxdim=maxdim, maxit=maxit,
                    ntol=ntol, sym=(g["sym_no"].tolist(), g["sym_nv"].tolist()), **kw)
        r.set_fock(lay.to_sym("FOO", s["FOO"].numpy()), lay.to_sym("FVO", s["FVO"].numpy()),
                   lay.to_sym("FVV", s["FVV"].numpy()), s["eps"])
        for c in CLASSES:
            r.set_ints(c, lay.to_sym(c, s[c].numpy()))
    else:
        r = RelCCSD(int(g["params"][0]), int(g["params"][1]), maxdim=maxdim, maxit=maxit,
                    ntol=ntol, **kw)
        r.set_system(s)
    return r, g


CASES = [("ccsd", "cfg1_6o40v"), ("ccsd", "conv_5o16v"), ("symccsd", "s8_8o24v"),
         ("symccsd", "s4_5o12v")]


@pytest.mark.parametrize("kind,name", CASES)
def test_graph_replays_equal_eager_iterations(kind, name):
    eager, g = make_solver(kind, name)
    graph, _ = make_solver(kind, name)
    graph.set_graph(True)
    n = int(g["nit"])
    eager.mp2()
    graph.mp2()
    ref = g["iterations"][1:]
    for i in range(n):
        a = eager.iterate()
        b = graph.iterate()     # iteration 1 eager, 2 captured, 3.. replays
        assert a == b, (i, a, b)
        assert abs(b[1] - ref[i][1]) < 1e-10
    ta, tb = eager.get_amplitudes(), graph.get_amplitudes()
    assert np.array_equal(ta[0], tb[0]) and np.array_equal(ta[1], tb[1])


def test_graph_replays_equal_eager_split_int8():
    eager, g = make_solver("ccsd", "conv_5o16v", precision=2)
    graph, _ = make_solver("ccsd", "conv_5o16v", precision=2)
    graph.set_graph(True)
    eager.mp2()
    graph.mp2()
    for i in range(8):
        assert eager.iterate() == graph.iterate(), i


def test_graph_replays_equal_eager_split_bf16():
    # the split-BF16 sequence: its slicing launches, the products that run
    # split in this mode (ladder, X, rings, AINTM, HVINTM, GV F1, tau AZ) and
    # the HINTM ring operand sliced straight from T2 and T1
    eager, g = make_solver("ccsd", "conv_5o16v", precision=1)
    graph, _ = make_solver("ccsd", "conv_5o16v", precision=1)
    graph.set_graph(True)
    eager.mp2()
    graph.mp2()
    ref = g["iterations"][1:]
    for i in range(min(8, int(g["nit"]))):
        a = eager.iterate()
        assert a == graph.iterate(), i
        assert abs(a[1] - ref[i][1]) < 1e-6 * abs(ref[i][1])


@pytest.mark.parametrize("mode", ["streams", "graph", "grouped"])
def test_batch_follows_each_reference_trajectory(mode):
    solvers, golds = [], []
    for kind, name in CASES + [("ccsd", "small_4o10v"), ("symccsd", "s2_4o12v")]:
        r, g = make_solver(kind, name)
        if mode == "graph":
            r.set_graph(True)
        r.mp2()
        solvers.append(r)
        golds.append(g)
    steps = max(int(g["nit"]) for g in golds)
    rows = [[] for _ in solvers]
    for _ in range(steps):
        for i, row in enumerate(RelCCSD.iterate_batch(solvers, grouped=(mode == "grouped"))):
            rows[i].append(row)
    for r, g, got in zip(solvers, golds, rows):
        ref = g["iterations"][1:]
        for (nit, e, rms), want in zip(got[:int(g["nit"])], ref):
            assert abs(e - want[1]) < 1e-10, (nit, e, want[1])
            assert abs(rms - want[2]) < 1e-9 * max(1.0, want[2])


def test_grouped_batch_equals_single_handle_iterations():
    # the grouped products are the same GEMM kernel over the same blocks:
    # a grouped batch of one system and a plain handle agree bit for bit
    for kind, name in (("symccsd", "s8_8o24v"), ("symccsd", "s4_5o12v")):
        a, g = make_solver(kind, name)
        b, _ = make_solver(kind, name)
        a.mp2()
        b.mp2()
        for _ in range(5):
            assert a.iterate() == RelCCSD.iterate_batch([b], grouped=True)[0]


@pytest.mark.parametrize("use_graph", [False, True])
def test_compiled_batch_of_symmetric_systems_follows_references(use_graph):
    # handles of one structure (all symmetric): the term sequence is recorded
    # once and replayed as one launch per position over all handles
    names_ = ["s2_4o12v", "s4_5o12v", "s8_8o24v"]
    solvers, golds, singles = [], [], []
    for name in names_:
        r, g = make_solver("symccsd", name)
        one, _ = make_solver("symccsd", name)
        if use_graph:
            r.set_graph(True)
        r.mp2()
        one.mp2()
        solvers.append(r)
        singles.append(one)
        golds.append(g)
    steps = max(int(g["nit"]) for g in golds)
    for it in range(steps):
        rows = RelCCSD.iterate_batch(solvers, grouped=True)
        for i, (r, g) in enumerate(zip(solvers, golds)):
            if it < int(g["nit"]):
                want = g["iterations"][1 + it]
                assert abs(rows[i][1] - want[1]) < 1e-10, (names_[i], it)
                assert abs(rows[i][2] - want[2]) < 1e-9 * max(1.0, want[2])
            # and bit for bit the single handle's eager iteration
            assert singles[i].iterate() == rows[i], (names_[i], it)
