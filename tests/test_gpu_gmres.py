"""The newton_krylov_solver's linear solve on the device: restarted GMRES
with left preconditioning (NLSolver_NK::GMRES).  Inexact Newton (linear
tolerance 1e-3) and tight solves reach the reference's converged
displacements; the linear solve matches the oracle's GMRES restatement."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

TOL_U = 1e-9


def make_solid(g):
    from tahoe_amd import api
    s = api.Solid(0)
    s.set_mesh(g["X"], g["conn"], g["mat"])
    for m, (E, nu) in enumerate(zip(g["E"], g["nu"])):
        s.set_material(m, *api.mu_kappa(E, nu))
    s.set_formulation(1 if "Q1P0" in str(g["element"]) else 0)
    s.set_kbc(g["kbc_nodes"], g["kbc_dofs"], g["kbc_values"])
    s.initialize()
    return s


@pytest.mark.parametrize("name", ["tl_n8", "q1p0_n4", "rve_n4"])
@pytest.mark.parametrize("lin_rtol,pc", [(1e-3, 0), (1e-10, 0), (1e-3, 2)])
def test_newton_krylov_reaches_reference(name, lin_rtol, pc):
    from tahoe_amd import api
    g = load_golden(name)
    s = make_solid(g)
    s.set_linear_solver(api.LIN_GMRES, 30)
    s.set_preconditioner(pc)
    lin = []
    us, iters = api.static_solve(s, int(g["num_steps"]), s.opts(lin_rtol=lin_rtol, lin_maxit=0),
                                 on_step=lambda k, st: lin.append(st.lin_iterations))
    print(f"# {name} rtol {lin_rtol} pc {pc}: Newton {iters} GMRES steps {lin}", flush=True)
    for k in range(us.shape[0]):
        ref = g["step_u"][k]
        assert np.abs(us[k] - ref).max() <= TOL_U * np.abs(ref).max(), (name, k)
    if lin_rtol <= 1e-10:
        assert list(iters) == g["step_iters"].tolist()
    s.close()


@pytest.mark.parametrize("restart", [30, 5])
def test_gmres_solve_matches_oracle(restart):
    """tg_solve with GMRES on the assembled tl_n4 stiffness: the solution of
    the oracle's GMRES to the same tolerance, and a small residual."""
    from oracle import oracle as O
    from tahoe_amd import api
    g = load_golden("tl_n4")
    s = make_solid(g)
    s.set_linear_solver(api.LIN_GMRES, restart)
    u = g["step_u"][0]
    s.set_displacement(u)
    s.form_lhs()
    val = s.msr_values()
    b = s.msr_bindx()
    rng = np.random.default_rng(12)
    rhs = rng.standard_normal(s.n_eq)
    x, st = s.solve(rhs, rtol=1e-10, maxit=0)
    xo, ok, steps = O.gmres_msr(b, val, rhs, restart, 1e-10)
    assert np.abs(x - xo).max() <= 1e-8 * np.abs(xo).max()
    r = rhs - O.msr_multx(b, val, x)
    d = val[:s.n_eq]
    assert np.linalg.norm(r / d) <= 1e-9 * np.linalg.norm(rhs / d)
    s.close()


def test_gmres_rejects_segments():
    from tahoe_amd import api, batch, rve
    pk = batch.pack([rve.Request(k, 3, api.TOTAL_LAGRANGIAN, 0.3, seed=5 + k).problem() for k in range(2)])
    s = api.Solid(0)
    s.set_mesh(pk.X, pk.conn, pk.mat)
    for m, (E, nu) in enumerate(pk.materials):
        s.set_material(m, *api.mu_kappa(E, nu))
    s.set_kbc(pk.kbc.nodes, pk.kbc.dofs, pk.kbc.values)
    s.set_segments(pk.node_begin)
    with pytest.raises(api.TahoeGpuError):
        s.set_linear_solver(api.LIN_GMRES, 30)
    s.close()
