"""large_strain_StVenant on the total-Lagrangian and Q1P0 kernels and nodal force
cards (the reference's beam decks): residual and stiffness against the
oracle's index-form restatement, both assembly paths and precisions, and
the reference's beam.Newton / beam.Newton.TotLag displacements."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

TOL_RK = 1e-12


def cube_solid(n=6, precision=0, seed=11, formulation=None):
    from tahoe_amd import api, mesh as M
    m = M.cube(n, seed=seed)
    kbc = M.kbc_uniaxial(m, 0.05)
    s = api.Solid(0)
    if formulation is not None:
        s.set_formulation(formulation)
    s.set_mesh(m.X, m.conn, m.mat)
    s.set_material(0, *api.mu_kappa(100.0, 0.25))
    s.set_material_model(api.MAT_STVENANT)
    s.set_precision(precision)
    s.set_kbc(kbc.nodes, kbc.dofs, kbc.values)
    s.initialize()
    return m, kbc, s


def test_residual_and_stiffness_match_oracle():
    from oracle import oracle as O
    m, kbc, s = cube_solid()
    eq, n = O.equation_numbers(m.nnd, kbc.nodes, kbc.dofs)
    rng = np.random.default_rng(4)
    u = 0.02 * rng.standard_normal((m.nnd, 3))
    s.set_displacement(u)
    R, _ = s.form_rhs()
    mu, ka = O.mu_kappa(100.0, 0.25)
    Ro = O.residual(O.TL_STVENANT, m.X, m.conn, m.mat, mu, ka, u, eq, n)
    assert np.abs(R - Ro).max() <= TOL_RK * np.abs(Ro).max()
    s.form_lhs()
    b = s.msr_bindx()
    vo = O.stiffness(O.TL_STVENANT, m.X, m.conn, m.mat, mu, ka, u, eq, n, b)
    assert np.abs(s.msr_values() - vo).max() <= TOL_RK * np.abs(vo).max()
    # the Simo material on the same state differs (the model switch is live)
    s2 = cube_solid()[2]
    s2.set_material_model(0)
    s2.initialize()
    s2.set_displacement(u)
    R2, _ = s2.form_rhs()
    assert np.abs(R2 - Ro).max() > 1e-6 * np.abs(Ro).max()
    s.close()
    s2.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_newton_both_assembly_paths(precision):
    """Newton's tangent straight into the operator or via the MSR: the same
    displacements (bitwise), converged to the oracle's."""
    from tahoe_amd import api
    out = []
    for mode in (0, 1):
        m, kbc, s = cube_solid(precision=precision)
        s.set_tuning(api.TUNE_NEWTON_ASSEMBLY, mode)
        us, iters = api.static_solve(s, 2, s.opts(lin_rtol=1e-12))
        out.append(us)
        s.close()
    assert np.array_equal(out[0], out[1])
    from oracle import oracle as O
    mu, ka = O.mu_kappa(100.0, 0.25)
    eq, n = O.equation_numbers(m.nnd, kbc.nodes, kbc.dofs)
    # oracle Newton for the same 2 load steps
    b = O.msr_pattern(n, m.conn, eq)
    u = np.zeros((m.nnd, 3))
    free = eq > 0
    for k in (1, 2):
        u[kbc.nodes, kbc.dofs] = kbc.values * k / 2
        R = O.residual(O.TL_STVENANT, m.X, m.conn, m.mat, mu, ka, u, eq, n)
        r0 = np.linalg.norm(R)
        for _ in range(20):
            val = O.stiffness(O.TL_STVENANT, m.X, m.conn, m.mat, mu, ka, u, eq, n, b)
            du, *_ = O.pcg_jacobi(b, val, R, rtol=1e-13)
            u[free] += du[eq[free] - 1]
            R = O.residual(O.TL_STVENANT, m.X, m.conn, m.mat, mu, ka, u, eq, n)
            if np.linalg.norm(R) < 1e-10 * r0:
                break
        assert np.abs(out[0][k - 1] - u).max() <= 1e-9 * np.abs(u).max(), k


@pytest.mark.parametrize("precision", [0, 1])
def test_beam_with_force_card(precision):
    """The reference's beam: the force card enters the residual; one step
    from rest reaches the reference's displacements (13 digits)."""
    from host_decks import read_geom
    from tahoe_amd import api
    X, conn, sets = read_geom(os.path.join(GOLDEN, "geometry", "beam.1.geom"))
    s = api.Solid(0)
    s.set_mesh(X, conn, np.zeros(len(conn), dtype=np.int32))
    s.set_material(0, *api.mu_kappa(100.0, 0.25))
    s.set_material_model(api.MAT_STVENANT)
    s.set_precision(precision)
    s.set_kbc(np.repeat(sets[1], 3), np.tile(np.arange(3), len(sets[1])), np.zeros(3 * len(sets[1])))
    s.set_forces(sets[2], np.zeros(len(sets[2]), dtype=np.int32), np.full(len(sets[2]), -0.1))
    s.initialize()
    s.apply_forces(1.0)
    us, iters = api.static_solve(s, 1, s.opts(rel_tol=1e-12, abs_tol=1e-12, lin_rtol=1e-13))
    ref = np.load(os.path.join(GOLDEN, "beam_newton_run.npz"))["u"][0]
    assert np.abs(us[0] - ref).max() <= 1e-10 * np.abs(ref).max()
    # forces scale with the schedule value: half the force, a different state
    s.reset_step()
    s.apply_forces(0.5)
    us2, _ = api.static_solve(s, 1, s.opts(rel_tol=1e-12, abs_tol=1e-12, lin_rtol=1e-13))
    assert 0.3 < np.abs(us2[0]).max() / np.abs(ref).max() < 0.7
    s.close()


def test_q1p0_residual_and_stiffness_match_oracle():
    """updated_lagrangian_Q1P0 with large_strain_StVenant (the reference's
    Q1P0.3D.a): the material sees F_bar.  R to 1e-12; the forward-difference
    tangent (eps = 1e-7 on both sides) to 6e-9 of max|K| (the forward
    difference's rounding noise, measured at ~3e-9 against the reference)."""
    from oracle import oracle as O
    from tahoe_amd import api
    m, kbc, s = cube_solid(n=4, formulation=api.Q1P0)
    eq, n = O.equation_numbers(m.nnd, kbc.nodes, kbc.dofs)
    rng = np.random.default_rng(5)
    u = 0.02 * rng.standard_normal((m.nnd, 3))
    s.set_displacement(u)
    R, _ = s.form_rhs()
    mu, ka = O.mu_kappa(100.0, 0.25)
    Ro = O.residual(O.Q1P0_STVENANT, m.X, m.conn, m.mat, mu, ka, u, eq, n)
    assert np.abs(R - Ro).max() <= TOL_RK * np.abs(Ro).max()
    s.form_lhs()
    b = s.msr_bindx()
    vo = O.stiffness(O.Q1P0_STVENANT, m.X, m.conn, m.mat, mu, ka, u, eq, n, b)
    assert np.abs(s.msr_values() - vo).max() <= 6e-9 * np.abs(vo).max()
    # the Simo Q1P0 element on the same state differs (the switch is live)
    Rs = O.residual(O.Q1P0, m.X, m.conn, m.mat, mu, ka, u, eq, n)
    assert np.abs(Rs - Ro).max() > 1e-6 * np.abs(Ro).max()
    s.close()


def test_q1p0_newton_both_precisions():
    """Q1P0 + St.Venant Newton, two uniaxial load steps: FP64 reaches the
    oracle's Newton (with the same forward-difference tangent) and the mixed
    path (FP32 operator, FP64 residual and Newton test) reaches FP64, both
    within 1e-9."""
    from oracle import oracle as O
    from tahoe_amd import api
    out = []
    for precision in (0, 1):
        m, kbc, s = cube_solid(n=4, precision=precision, formulation=api.Q1P0)
        us, iters = api.static_solve(s, 2, s.opts(lin_rtol=1e-12))
        out.append(us)
        s.close()
    mu, ka = O.mu_kappa(100.0, 0.25)
    eq, n = O.equation_numbers(m.nnd, kbc.nodes, kbc.dofs)
    b = O.msr_pattern(n, m.conn, eq)
    u = np.zeros((m.nnd, 3))
    free = eq > 0
    for k in (1, 2):
        u[kbc.nodes, kbc.dofs] = kbc.values * k / 2
        R = O.residual(O.Q1P0_STVENANT, m.X, m.conn, m.mat, mu, ka, u, eq, n)
        r0 = np.linalg.norm(R)
        for _ in range(20):
            val = O.stiffness(O.Q1P0_STVENANT, m.X, m.conn, m.mat, mu, ka, u, eq, n, b)
            du, *_ = O.pcg_jacobi(b, val, R, rtol=1e-13)
            u[free] += du[eq[free] - 1]
            R = O.residual(O.Q1P0_STVENANT, m.X, m.conn, m.mat, mu, ka, u, eq, n)
            if np.linalg.norm(R) < 1e-10 * r0:
                break
        scale = np.abs(u).max()
        assert np.abs(out[0][k - 1] - u).max() <= 1e-9 * scale, k
        assert np.abs(out[1][k - 1] - out[0][k - 1]).max() <= 1e-9 * scale, k
    # the mixed leg ran its own (FP32-operator) arithmetic
    assert not np.array_equal(out[1], out[0])
