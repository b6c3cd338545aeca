"""The equation-system structure built on the device (node graph, MSR
pattern) for connectivity outside the structured meshes of the other tests:
nodes shared by more than 8 elements take the sorted-insertion path of the
neighbour kernel, an isolated node has no neighbours, and a node shared by
more than 64 elements is refused.  Only the structure is checked (the hub
elements' geometry is arbitrary), bit-exactly against the oracle's
MSRBuilderT restatement."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def hub_mesh(fans, n=3, seed=5):
    """An n^3 cube, then for each fan size f a hub node shared by f elements
    whose other 7 nodes are consecutive on a ring of 3 f + 4 new nodes, then
    one node no element uses."""
    from tahoe_amd import mesh as M
    m = M.cube(n)
    X, conn = [m.X], [m.conn]
    nnd = m.X.shape[0]
    rng = np.random.default_rng(seed)
    for f in fans:
        hub, ring = nnd, 3 * f + 4
        el = [[hub] + [hub + 1 + (3 * e + k) % ring for k in range(7)] for e in range(f)]
        conn.append(np.array(el, dtype=np.int32))
        X.append(rng.random((1 + ring, 3)))
        nnd += 1 + ring
    X.append(rng.random((1, 3)))
    return np.concatenate(X), np.concatenate(conn).astype(np.int32)


def set_up(X, conn, kbc_every=5):
    from tahoe_amd import api
    s = api.Solid(0)
    s.set_mesh(X, conn, np.zeros(conn.shape[0], dtype=np.int32))
    s.set_material(0, *api.mu_kappa(1000.0, 0.3))
    dofs = np.arange(0, 3 * X.shape[0], kbc_every)
    s.set_kbc((dofs // 3).astype(np.int32), (dofs % 3).astype(np.int32), np.zeros(len(dofs)))
    return s, dofs


@pytest.mark.parametrize("fans", [(12,), (9, 40, 64)])
def test_pattern_with_high_valence_nodes(fans):
    from oracle import oracle as O
    X, conn = hub_mesh(fans)
    s, dofs = set_up(X, conn)
    try:
        s.initialize()
        eq, n = O.equation_numbers(X.shape[0], dofs // 3, dofs % 3)
        assert s.n_eq == n and np.array_equal(s.eqnos(), eq)
        assert np.array_equal(s.msr_bindx(), O.msr_pattern(n, conn, eq))
    finally:
        s.close()


def test_node_of_more_than_64_elements_is_refused_and_context_recovers():
    from tahoe_amd import api
    from oracle import oracle as O
    X, conn = hub_mesh((65,))
    s, _ = set_up(X, conn)
    try:
        with pytest.raises(api.TahoeGpuError) as e:
            s.initialize()
        assert e.value.code == 4  # TG_OUT_OF_RANGE
        X2, conn2 = hub_mesh((20,))
        s.set_mesh(X2, conn2, np.zeros(conn2.shape[0], dtype=np.int32))
        dofs = np.arange(0, 3 * X2.shape[0], 5)
        s.set_kbc((dofs // 3).astype(np.int32), (dofs % 3).astype(np.int32), np.zeros(len(dofs)))
        s.initialize()
        eq, n = O.equation_numbers(X2.shape[0], dofs // 3, dofs % 3)
        assert np.array_equal(s.msr_bindx(), O.msr_pattern(n, conn2, eq))
    finally:
        s.close()
