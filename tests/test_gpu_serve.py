"""Continuous-batching server: requests of both formulations arriving over
time are grouped into mini-batches on several contexts at once; every
result equals the same request solved alone (bitwise: a sub-problem's
arithmetic does not depend on the batch it lands in; a batch that fails is
re-solved request by request through the same single-request path) and
matches the plain FP64 solve within 1e-9."""
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
    by_id = {r.rid: r for r in reqs}
    assert all(x.status == 0 for x in st.results.values()), \
        [(x.rid, by_id[x.rid].L, by_id[x.rid].formulation, x.status, x.load_steps, x.batch_size)
         for x in st.results.values() if x.status != 0]
    assert max(b[1] for b in st.batches) > 1, "no request was batched with another"
    fp64 = api.Solid.opts(lin_rtol=1e-10)
    for r in reqs:
        got = st.results[r.rid].u
        alone, _, _ = serve.solve_request(r, api.MIXED, srv.opts)
        assert np.array_equal(got, alone), r.rid
        ref, _, _ = serve.solve_request(r, api.FP64, fp64)
        assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max(), r.rid
    lat = serve.latency_summary(st)
    assert lat["p99_s"] >= lat["p50_s"] > 0


def test_large_near_incompressible_request_on_multigrid():
    """A large Q1P0 request (edge >= MULTIGRID_MIN_EDGE) is served on its own
    context with the multigrid preconditioner while small requests are
    batched by size class; it equals itself solved alone (bitwise) and the
    plain FP64 Jacobi solve within 1e-9."""
    from tahoe_amd import api, rve, serve
    big = rve.Request(100, serve.MULTIGRID_MIN_EDGE, api.Q1P0, 0.499, seed=4321, arrival=0.0, isochoric=True)
    assert serve.uses_multigrid(big)
    small = rve.stream_requests(8, rate=400.0, lo=3, hi=6, seed=17)
    reqs = sorted(small + [big], key=lambda r: r.arrival)
    srv = serve.RVEServer(precision=api.MIXED, workers=2, max_batch=4)
    st = srv.run(reqs, realtime=True)
    assert all(x.status == 0 for x in st.results.values())
    assert st.results[big.rid].batch_size == 1
    got = st.results[big.rid].u
    alone, _, _ = serve.solve_request(big, api.MIXED, srv.opts)
    assert np.array_equal(got, alone)
    ref, _, _ = serve.solve_request(big, api.FP64, api.Solid.opts(lin_rtol=1e-10), multigrid=False)
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()
