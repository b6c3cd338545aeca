"""Batched independent problems in one context (segments): every problem's
result equals its reference fixture (converged u within 1e-9, same Newton
iteration counts) and is bitwise identical alone and inside the batch."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def golden_problem(name):
    from tahoe_amd import batch, mesh as M
    g = load_golden(name)
    m = M.HexMesh(g["X"], g["conn"], g["mat"], {}, None)
    k = M.KBC(g["kbc_nodes"], g["kbc_dofs"], g["kbc_values"])
    return g, batch.Problem(m, k, list(g["E"]), list(g["nu"]))


def rve_problem(n, seed):
    from tahoe_amd import batch, mesh as M
    m, H = M.rve(n, seed)
    return batch.Problem(m, M.kbc_affine(m, H), [1000.0, 10000.0], [0.3, 0.3])


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_matches_fixtures_and_alone(precision):
    from tahoe_amd import api, batch
    names = ["tl_n2", "rve_n3", "rve_n4"]
    gs, probs = zip(*[golden_problem(n) for n in names])
    probs = list(probs) + [rve_problem(5, 77), rve_problem(6, 78)]
    opts = api.Solid.opts(lin_rtol=1e-12 if precision == 0 else 1e-6)
    pk = batch.pack(probs)
    s = batch.make_solid(pk, precision=precision)
    us, stats = batch.static_solve_batch(s, pk, 2, opts)
    for i, g in enumerate(gs):
        ref = g["step_u"][-1]
        err = np.abs(us[i] - ref).max() / np.abs(ref).max()
        assert err <= 1e-9, (names[i], err)
        if precision == 0:
            assert [st[i].iterations for st in stats] == g["step_iters"].tolist()
    for i, p in enumerate(probs):
        pk1 = batch.pack([p])
        s1 = batch.make_solid(pk1, precision=precision)
        u1, st1 = batch.static_solve_batch(s1, pk1, 2, opts)
        assert np.array_equal(u1[0], us[i]), i
        assert [st[0].iterations for st in st1] == [st[i].iterations for st in stats]


def test_single_segment_batch_equals_newton():
    """A one-problem batch runs the same arithmetic as tg_newton."""
    from tahoe_amd import api, batch
    g, p = golden_problem("tl_n4")
    pk = batch.pack([p])
    s = batch.make_solid(pk)
    s.apply_kbc(0.2)
    st = s.newton_batch()
    ub = s.displacement()
    s2 = batch.make_solid(pk)
    s2.apply_kbc(0.2)
    st2 = s2.newton()
    assert np.array_equal(ub, s2.displacement())
    assert st[0].iterations == st2.iterations


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replayed_pcg_equals_direct_launches(precision):
    """The PCG chunk captured once as a HIP graph and replayed across every
    Newton iteration and load step (with eager assembly kernels and host
    syncs in between) gives bitwise the direct-launch results."""
    from tahoe_amd import api, batch
    probs = [golden_problem(n)[1] for n in ["tl_n2", "rve_n3", "rve_n4"]] + [rve_problem(5, 81)]
    opts = api.Solid.opts(lin_rtol=1e-12 if precision == 0 else 1e-6)
    out = []
    for graph in (0, 1):
        pk = batch.pack(probs)
        s = batch.make_solid(pk, precision=precision)
        s.set_tuning(api.TUNE_PCG_GRAPH, graph)
        us, st = batch.static_solve_batch(s, pk, 2, opts)
        out.append((us, [[x.iterations for x in row] for row in st]))
        s.close()
    assert out[0][1] == out[1][1]
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)


def test_default_pcg_variant_by_size():
    """Untuned PCG variant: a single problem of 2^17 nodes or more reads the
    lower blocks through their partners with the direction in its own pass
    (7, unfused); a batch of that size reads the full operator with the
    unrolled loop, direction in its own pass (2, unfused); anything smaller
    fuses the direction into the hardware-order SpMV (0, fused)."""
    from tahoe_amd import api, batch
    small = batch.pack([rve_problem(4, 1), rve_problem(5, 2)])
    big = batch.pack([rve_problem(14, 10 + k) for k in range(40)])
    assert big.X.shape[0] >= 1 << 17
    single = batch.pack([rve_problem(50, 3)])
    assert single.X.shape[0] >= 1 << 17
    for pk, want in ((small, (0, 1)), (big, (2, 0)), (single, (7, 0))):
        s = batch.make_solid(pk, precision=api.MIXED)
        try:
            assert (s.tuning(api.TUNE_SPMV_XCD), s.tuning(api.TUNE_PCG_FUSED)) == want
        finally:
            s.close()
