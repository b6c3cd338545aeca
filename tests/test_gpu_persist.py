"""Persistent PCG (one workgroup runs a small segment's whole iteration):
the iterates are bitwise those of the kernel-per-step iteration -- the same
sums in the same order -- for a single problem and for a batch whose
segments fall on both sides of the size threshold, in FP64 and mixed
precision, with point and block Jacobi."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve(pk, formulation, precision, pc, persist, fused):
    from tahoe_amd import api, batch
    s = batch.make_solid(pk, formulation=formulation, precision=precision, precond=pc)
    try:
        s.set_tuning(api.TUNE_PCG_PERSIST, persist)
        s.set_tuning(api.TUNE_PCG_FUSED, fused)
        assert s.tuning(api.TUNE_PCG_PERSIST) == persist
        us, st = batch.solve_one_step_partial(s, pk, api.Solid.opts(lin_rtol=1e-6))
        return us, [(x.status, x.iterations, x.lin_iterations) for x in st]
    finally:
        s.close()


@pytest.mark.parametrize("form,pc", [("Q1", 0), ("Q1", 1), ("Q1P0", 1)])
@pytest.mark.parametrize("precision", [0, 1])
def test_persistent_pcg_bitwise(form, pc, precision):
    from tahoe_amd import api, batch, rve
    f = api.Q1P0 if form == "Q1P0" else api.TOTAL_LAGRANGIAN
    nu = 0.499 if form == "Q1P0" else 0.3
    prec = api.MIXED if precision else api.FP64
    pcs = (api.PC_JACOBI, api.PC_BLOCK_JACOBI)[pc]
    # edges 3, 4, 6 and 9: 1, 2, 6 and 16 slices of 64 nodes (the persistent
    # kernel takes segments of up to 16)
    probs = [rve.Request(i, L, f, nu, 500 + i, isochoric=form == "Q1P0").problem() for i, L in enumerate([3, 9, 4, 6])]
    for pk in (batch.pack(probs), batch.pack(probs[1:2])):
        ref = _solve(pk, f, prec, pcs, 0, 0)
        assert all(x[0] == 0 for x in ref[1])
        for persist, fused in ((4, 0), (16, 1), (1, 1)):
            got = _solve(pk, f, prec, pcs, persist, fused)
            assert got[1] == ref[1], (persist, fused)
            for a, b in zip(got[0], ref[0]):
                assert np.array_equal(a, b), (persist, fused)


def test_persistent_threshold_set_after_initialize():
    """Changing the threshold on an initialised context takes effect at the
    next solve; the result stays bitwise the same."""
    from tahoe_amd import api, batch, rve
    probs = [rve.Request(i, L, api.TOTAL_LAGRANGIAN, 0.3, 900 + i).problem() for i, L in enumerate([5, 7])]
    pk = batch.pack(probs)
    s = batch.make_solid(pk, precision=api.MIXED)
    try:
        out = []
        for persist in (0, 16, 0):
            s.set_tuning(api.TUNE_PCG_PERSIST, persist)
            us, st = batch.solve_one_step_partial(s, pk, api.Solid.opts(lin_rtol=1e-6))
            out.append((us, [x.lin_iterations for x in st]))
        for o in out[1:]:
            assert o[1] == out[0][1]
            assert all(np.array_equal(a, b) for a, b in zip(o[0], out[0][0]))
    finally:
        s.close()
