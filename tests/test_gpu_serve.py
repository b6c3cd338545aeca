"""Continuous-batching server: requests of both formulations arriving over
time are grouped into mini-batches on several contexts at once; every
result equals the same request solved alone (bitwise: a sub-problem's
arithmetic does not depend on the batch it lands in) and matches the plain
FP64 solve within 1e-9."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_served_equals_alone():
    from tahoe_amd import api, batch, rve, serve
    reqs = rve.stream_requests(24, rate=400.0, lo=3, hi=7, seed=91)
    assert {r.formulation for r in reqs} == {api.TOTAL_LAGRANGIAN, api.Q1P0}
    srv = serve.RVEServer(precision=api.MIXED, workers=3, max_batch=6)
    st = srv.run(reqs, realtime=True)
    assert len(st.results) == len(reqs)
    assert all(x.status == 0 for x in st.results.values()), [(x.rid, x.status) for x in st.results.values()]
    assert max(b[1] for b in st.batches) > 1, "no request was batched with another"
    fp64 = api.Solid.opts(lin_rtol=1e-10)
    for r in reqs:
        got = st.results[r.rid].u
        steps = st.results[r.rid].load_steps
        pk = batch.pack([r.problem()])
        s = batch.make_solid(pk, formulation=r.formulation, precision=api.MIXED)
        alone, _ = batch.static_solve_batch(s, pk, steps, srv.opts)
        s.close()
        assert np.array_equal(got, alone[0]), r.rid
        s = batch.make_solid(pk, formulation=r.formulation, precision=api.FP64)
        ref, _ = batch.static_solve_batch(s, pk, steps, fp64)
        s.close()
        assert np.abs(got - ref[0]).max() <= 1e-9 * np.abs(ref[0]).max(), r.rid
    lat = serve.latency_summary(st)
    assert lat["p99_s"] >= lat["p50_s"] > 0
