"""The 128^3 mesh of the bench line (2,097,152 elements, 2,146,689 nodes):
residual and stiffness against the oracle's coloured passes on all host
threads (FP64 bar, 1e-12 of the largest entry), the MSR pattern and the
equation numbers exactly, at a non-trivial state; and the size-independent
property that every PCG variant (SpMV forms, fused direction) gives
bitwise the same iterates at this size."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _threads():
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = os.cpu_count() or 1
    return max(1, min(n, int(env))) if env.isdigit() and int(env) > 0 else n


@pytest.fixture(scope="module")
def cube128():
    from tahoe_amd import api, mesh as M
    m = M.cube(128, seed=2024)
    kbc = M.kbc_uniaxial(m, 0.05)
    mu, ka = api.mu_kappa(1000.0, 0.3)
    s = api.Solid(0)
    s.set_mesh(m.X, m.conn, m.mat)
    s.set_material(0, mu, ka)
    s.set_kbc(kbc.nodes, kbc.dofs, kbc.values)
    n, _ = s.initialize()
    yield m, kbc, s, n, (mu, ka)
    s.close()


def test_128cube_residual_stiffness_pattern_against_oracle(cube128):
    from oracle import oracle as O
    m, kbc, s, n, (mu, ka) = cube128
    eq, n_o = O.equation_numbers(m.nnd, kbc.nodes, kbc.dofs)
    assert n == n_o and np.array_equal(s.eqnos(), eq)
    rng = np.random.default_rng(128)
    u = 2e-4 * rng.standard_normal((m.nnd, 3))
    s.set_displacement(u)
    R, _ = s.form_rhs()
    s.form_lhs()
    b = s.msr_bindx()
    assert b[0] == n + 1 and len(b) == s.msr_len
    val = s.msr_values()
    Ro, vo = O.element_passes_parallel(O.TL, m.X, m.conn, m.mat, mu, ka, u, eq, n, b, O.color_elements(m.nnd, m.conn),
                                       _threads())
    assert np.abs(R - Ro).max() <= 1e-12 * np.abs(Ro).max()
    assert np.abs(val - vo).max() <= 1e-12 * np.abs(vo).max()
    # the whole upper MSR pattern (row pointers and columns) is the oracle's
    assert np.array_equal(b, O.msr_pattern(n, m.conn, eq))


def test_128cube_pcg_variants_bitwise(cube128):
    from tahoe_amd import api
    m, kbc, s, n, _ = cube128
    out = []
    for fused, v in ((0, 0), (0, 7), (1, 0), (1, 5), (0, 3)):
        s.set_tuning(api.TUNE_PCG_FUSED, fused)
        s.set_tuning(api.TUNE_SPMV_XCD, v)
        s.reset_step()
        s.apply_kbc(0.2)
        st = s.newton_correction(s.opts(lin_rtol=1e-10, lin_maxit=64))
        out.append((s.displacement(), st.lin_iterations))
    s.set_tuning(api.TUNE_PCG_FUSED, -1)
    s.set_tuning(api.TUNE_SPMV_XCD, -1)
    assert out[0][1] == 64
    for o in out[1:]:
        assert o[1] == out[0][1]
        assert np.array_equal(o[0], out[0][0])
