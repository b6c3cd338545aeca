"""GPU parity of the total-Lagrangian path through the C ABI.

Reference fixtures (tests/golden/*.npz, written by the reference CPU build):
equation numbers and the MSR pattern must match bit-exactly; the residual
and the assembled MSR stiffness at every dumped Newton iterate within 1e-12
relative (to the largest entry of the iteration-0 vector/matrix); converged
displacements within 1e-9 relative.  Larger meshes are checked against the
C oracle and through size-independent properties.
"""
import numpy as np
import pytest

from conftest import golden_K_dense, golden_R_eq, load_golden, msr_dense, msr_pattern_dense

pytestmark = pytest.mark.gpu

TOL_RK = 1e-12
TOL_U = 1e-9


def make_solid(g, formulation=None):
    from tahoe_amd import api
    s = api.Solid(0)
    s.set_mesh(g["X"], g["conn"], g["mat"])
    for m, (E, nu) in enumerate(zip(g["E"], g["nu"])):
        s.set_material(m, *api.mu_kappa(E, nu))
    s.set_formulation(formulation if formulation is not None else (1 if "Q1P0" in str(g["element"]) else 0))
    s.set_kbc(g["kbc_nodes"], g["kbc_dofs"], g["kbc_values"])
    s.initialize()
    return s


TL_CASES = ["tl_n2", "tl_n4", "rve_n3", "rve_n4"]


@pytest.mark.parametrize("name", TL_CASES)
def test_eqnos_and_pattern(name):
    from oracle import oracle as O
    g = load_golden(name)
    s = make_solid(g)
    assert np.array_equal(s.eqnos(), g["eqnos"])
    n = s.n_eq
    b = s.msr_bindx()
    ob = O.msr_pattern(n, g["conn"], g["eqnos"])
    assert np.array_equal(b, ob)
    P = msr_pattern_dense(b, n)
    G = golden_K_dense(g, 0, n) != 0
    assert not np.any(G & ~P)
    if not name.startswith("rve"):
        # perturbed meshes: every pattern entry is non-zero in the reference dump
        assert np.array_equal(P, G)


@pytest.mark.parametrize("name", TL_CASES)
def test_residual_and_stiffness_at_reference_iterates(name):
    g = load_golden(name)
    s = make_solid(g)
    n = s.n_eq
    b = s.msr_bindx()
    R0 = np.abs(golden_R_eq(g, 0)).max()
    nsolve = len(g["lhs_ptr"]) - 1
    K0 = np.abs(golden_K_dense(g, 0, n)).max()
    for k in range(g["it_u"].shape[0]):
        s.set_displacement(g["it_u"][k])
        R, nrm = s.form_rhs()
        Rg = golden_R_eq(g, k)
        assert np.abs(R - Rg).max() <= TOL_RK * R0, (name, k)
        assert abs(nrm - np.linalg.norm(R)) <= 1e-12 * max(nrm, 1e-300)
        if k < nsolve:
            s.form_lhs()
            A = msr_dense(b, s.msr_values(), n)
            Gk = golden_K_dense(g, k, n)
            assert np.abs(A - Gk).max() <= TOL_RK * K0, (name, k)


@pytest.mark.parametrize("name", ["tl_n2", "tl_n4", "tl_n8", "tl_n4_s20", "rve_n3", "rve_n4", "tl_n16"])
def test_converged_displacements(name):
    from tahoe_amd import api
    g = load_golden(name)
    s = make_solid(g)
    step_u, iters = api.static_solve(s, int(g["num_steps"]))
    for k in range(step_u.shape[0]):
        ref = g["step_u"][k]
        err = np.abs(step_u[k] - ref).max() / np.abs(ref).max()
        assert err <= TOL_U, (name, k, err)
    assert list(iters) == g["step_iters"].tolist()


def test_multx_copy_diagonal_and_solve_match_oracle():
    from oracle import oracle as O
    g = load_golden("tl_n8")
    s = make_solid(g)
    n = s.n_eq
    u = g["step_u"][0]
    s.set_displacement(u)
    s.form_lhs()
    val = s.msr_values()
    b = s.msr_bindx()
    mu, ka = O.mu_kappa(1000.0, 0.3)
    oval = O.stiffness(O.TL, g["X"], g["conn"], g["mat"], mu, ka, u, g["eqnos"], n, b)
    assert np.abs(val - oval).max() <= TOL_RK * np.abs(oval).max()
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n)
    y = s.multx(x)
    yo = O.msr_multx(b, oval, x)
    assert np.abs(y - yo).max() <= 1e-12 * np.abs(yo).max()
    assert np.array_equal(s.copy_diagonal(), val[:n])
    rhs = rng.standard_normal(n)
    xs, st = s.solve(rhs, rtol=1e-12)
    assert st.converged
    r = rhs - O.msr_multx(b, oval, xs)
    assert np.linalg.norm(r) <= 1e-11 * np.linalg.norm(rhs)


@pytest.mark.parametrize("name", ["tl_n8", "tl_n16"])
def test_pcg_iteration_count_matches_oracle(name):
    """Jacobi-PCG on a Newton system (K and R at a dumped iterate) takes the
    oracle's iteration count to 1e-10: the same algorithm, so only rounding
    may move the count (by one iteration or 1 %)."""
    from oracle import oracle as O
    g = load_golden(name)
    s = make_solid(g)
    u = g["step_u"][0]
    s.set_displacement(u)
    s.form_lhs()
    R, _ = s.form_rhs()
    val, b = s.msr_values(), s.msr_bindx()
    xs, st = s.solve(R, rtol=1e-10)
    xo, its, rr, rc = O.pcg_jacobi(b, val, R, rtol=1e-10)
    assert rc == 0 and st.converged
    assert abs(st.iterations - its) <= max(1, its // 100), (st.iterations, its)
    assert np.abs(xs - xo).max() <= 1e-8 * np.abs(xo).max()


def test_bad_jacobian_is_reported():
    """An inverted element gives status 7 (kBadJacobianDet), as
    TotalLagrangianT::FormKd throws (TotalLagrangianT.cpp:127-130)."""
    from tahoe_amd import api
    g = load_golden("tl_n2")
    s = make_solid(g)
    u = np.zeros((s.nnd, 3))
    u[13, 2] = -0.9  # push the centre node through the bottom layer
    s.set_displacement(u)
    with pytest.raises(api.TahoeGpuError) as e:
        s.form_rhs()
    assert e.value.code == 7
    # the handle stays usable
    s.set_displacement(np.zeros((s.nnd, 3)))
    R, nrm = s.form_rhs()
    assert nrm == 0.0


def test_large_mesh_against_oracle_and_invariants():
    """32^3 mesh: residual/stiffness against the oracle at a non-trivial state;
    equilibrium of rigid translation (R independent of a uniform shift of u)
    and symmetry of K through Multx."""
    from oracle import oracle as O
    from tahoe_amd import api, mesh as M
    m = M.cube(32, seed=7)
    kbc = M.kbc_uniaxial(m, 0.05)
    s = api.Solid(0)
    s.set_mesh(m.X, m.conn, m.mat)
    mu, ka = api.mu_kappa(1000.0, 0.3)
    s.set_material(0, mu, ka)
    s.set_kbc(kbc.nodes, kbc.dofs, kbc.values)
    n, L = s.initialize()
    eq, n2 = O.equation_numbers(m.nnd, kbc.nodes, kbc.dofs)
    assert n == n2 and np.array_equal(s.eqnos(), eq)
    rng = np.random.default_rng(3)
    u = 1e-3 * rng.standard_normal((m.nnd, 3))
    s.set_displacement(u)
    R, _ = s.form_rhs()
    Ro = O.residual(O.TL, m.X, m.conn, m.mat, mu, ka, u, eq, n)
    assert np.abs(R - Ro).max() <= 1e-12 * np.abs(Ro).max()
    s.form_lhs()
    val = s.msr_values()
    b = s.msr_bindx()
    assert np.array_equal(b, O.msr_pattern(n, m.conn, eq))
    vo = O.stiffness(O.TL, m.X, m.conn, m.mat, mu, ka, u, eq, n, b)
    assert np.abs(val - vo).max() <= 1e-12 * np.abs(vo).max()
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    assert abs(y @ s.multx(x) - x @ s.multx(y)) <= 1e-10 * np.linalg.norm(x) * np.linalg.norm(y) * np.abs(vo).max()


@pytest.mark.parametrize("shape", list(range(19)))
@pytest.mark.parametrize("chunk", [0, 100])
def test_stiffness_every_launch_shape(shape, chunk):
    """Every compiled workgroup shape and launch-range split assembles the
    same stiffness (the oracle's, to the FP64 bar) on a ragged element count
    (tl_n8: 512 elements, not a multiple of any shape; chunk 100 splits every
    colour into several launches)."""
    from oracle import oracle as O
    from tahoe_amd import api
    g = load_golden("tl_n8")
    s = make_solid(g)
    assert len(api.STIFF_SHAPE_EPB) == 19
    with pytest.raises(api.TahoeGpuError):
        s.set_tuning(api.TUNE_STIFF_SHAPE, len(api.STIFF_SHAPE_EPB))
    s.set_tuning(api.TUNE_STIFF_SHAPE, shape)
    s.set_tuning(api.TUNE_STIFF_CHUNK, chunk)
    assert s.tuning(api.TUNE_STIFF_SHAPE) == shape and s.tuning(api.TUNE_STIFF_CHUNK) == chunk
    u = g["step_u"][0]
    s.set_displacement(u)
    s.form_lhs()
    val = s.msr_values()
    mu, ka = O.mu_kappa(1000.0, 0.3)
    oval = O.stiffness(O.TL, g["X"], g["conn"], g["mat"], mu, ka, u, g["eqnos"], s.n_eq, s.msr_bindx())
    assert np.abs(val - oval).max() <= TOL_RK * np.abs(oval).max()


@pytest.mark.parametrize("name", ["tl_n8", "rve_n4"])
@pytest.mark.parametrize("precision", [0, 1])
def test_operator_assembly_equals_msr_expansion(name, precision):
    """Newton's tangent assembled straight into the solver operator gives
    the operator that expanding the assembled MSR gives (same sums, same
    order): the Newton corrections agree bitwise, prescribed dofs included.
    Both paths run the same kernel (shape 9; the matrix-core shapes assemble
    the MSR only)."""
    from tahoe_amd import api
    g = load_golden(name)
    out = []
    for mode in (0, 1):
        s = make_solid(g)
        if precision:
            s.set_precision(api.MIXED)
            s.initialize()
        s.set_tuning(api.TUNE_STIFF_SHAPE, 9)
        s.set_tuning(api.TUNE_NEWTON_ASSEMBLY, mode)
        s.apply_kbc(1.0 / int(g["num_steps"]))
        st = s.newton_correction(s.opts(lin_rtol=1e-10))
        out.append((s.displacement(), st.lin_iterations))
    for o in out[1:]:
        assert o[1] == out[0][1]
        assert np.array_equal(o[0], out[0][0])


@pytest.mark.parametrize("name", ["tl_n8", "rve_n4", "q1p0_n4"])
def test_block_jacobi_newton(name):
    """The 3x3 nodal block-Jacobi preconditioner converges to the reference's
    displacements (1e-9) with the same Newton counts and fewer PCG
    iterations than point Jacobi."""
    from tahoe_amd import api
    g = load_golden(name)
    its = {}
    for pc in (api.PC_JACOBI, api.PC_BLOCK_JACOBI):
        s = make_solid(g)
        s.set_preconditioner(pc)
        lin = []
        us, iters = api.static_solve(s, int(g["num_steps"]), s.opts(lin_rtol=1e-12),
                                     on_step=lambda k, st: lin.append(st.lin_iterations))
        for k in range(us.shape[0]):
            ref = g["step_u"][k]
            assert np.abs(us[k] - ref).max() <= TOL_U * np.abs(ref).max(), (name, pc, k)
        assert list(iters) == g["step_iters"].tolist()
        its[pc] = sum(lin)
    assert its[api.PC_BLOCK_JACOBI] < its[api.PC_JACOBI], its


@pytest.mark.parametrize("precision", [0, 1])
def test_spmv_xcd_mapping_bitwise(precision):
    """The SpMV variants change only which workgroup computes which rows,
    how the inner loop is issued (padding slots add exact zeros) and where a
    lower block's values are read from (its own slot, or its upper partner's
    transposed: the same values), with the direction formed inside the SpMV
    or in its own pass: the Newton correction is bitwise equal for all
    sixteen."""
    from tahoe_amd import api
    g = load_golden("rve_n4")
    out = []
    for fused, xcd in [(f, v) for f in (1, 0) for v in range(8)]:
        s = make_solid(g)
        if precision:
            s.set_precision(api.MIXED)
            s.initialize()
        s.set_tuning(api.TUNE_PCG_FUSED, fused)
        s.set_tuning(api.TUNE_SPMV_XCD, xcd)
        s.apply_kbc(1.0)
        st = s.newton_correction(s.opts(lin_rtol=1e-10))
        out.append((s.displacement(), st.lin_iterations))
    for o in out[1:]:
        assert o[1] == out[0][1]
        assert np.array_equal(o[0], out[0][0])


@pytest.mark.parametrize("reuse", [0, 6])
@pytest.mark.parametrize("name,precision", [("tl_n8", 0), ("tl_n16", 0), ("rve_n4", 0), ("q1p0_n4", 0),
                                            ("tl_n16", 1)])
def test_amg_newton(name, precision, reuse):
    """The aggregation-multigrid preconditioner changes only the inner
    solve: the static solve reaches the reference's displacements (1e-9)
    with the reference's Newton counts, in far fewer PCG iterations than
    point Jacobi on the larger meshes -- with a numeric set-up of the coarse
    levels for every tangent (reuse 0) and with one set-up per Newton solve
    serving its later tangents (reuse 6, the default)."""
    from tahoe_amd import api
    g = load_golden(name)
    its = {}
    for pc in (api.PC_JACOBI, api.PC_AMG):
        s = make_solid(g)
        if precision:
            s.set_precision(api.MIXED)
            s.initialize()
        s.set_preconditioner(pc)
        s.set_tuning(api.TUNE_AMG_REUSE, reuse)
        assert s.tuning(api.TUNE_AMG_REUSE) == reuse
        lin = []
        us, iters = api.static_solve(s, int(g["num_steps"]), s.opts(lin_rtol=1e-12),
                                     on_step=lambda k, st: lin.append(st.lin_iterations))
        print(f"# {name} precision {precision} pc {pc}: PCG iterations per step {lin}", flush=True)
        for k in range(us.shape[0]):
            ref = g["step_u"][k]
            assert np.abs(us[k] - ref).max() <= TOL_U * np.abs(ref).max(), (name, pc, k)
        assert list(iters) == g["step_iters"].tolist()
        its[pc] = sum(lin)
        s.close()
    if g["X"].shape[0] > 500:
        assert its[api.PC_AMG] * 3 < its[api.PC_JACOBI], its


def test_amg_set_after_initialize_and_rejects_segments():
    """The multigrid structure is built when the preconditioner is chosen
    after tg_initialize; a batch of segments is refused."""
    from tahoe_amd import api
    g = load_golden("tl_n8")
    s = make_solid(g)
    s.set_preconditioner(api.PC_AMG)
    s.apply_kbc(1.0 / int(g["num_steps"]))
    st = s.newton_correction(s.opts(lin_rtol=1e-10))
    assert st.lin_iterations > 0
    s.close()
    from tahoe_amd import batch, rve
    pk = batch.pack([rve.Request(k, 3, api.TOTAL_LAGRANGIAN, 0.3, seed=5 + k).problem() for k in range(2)])
    s = api.Solid(0)
    s.set_mesh(pk.X, pk.conn, pk.mat)
    for m, (E, nu) in enumerate(pk.materials):
        s.set_material(m, *api.mu_kappa(E, nu))
    s.set_kbc(pk.kbc.nodes, pk.kbc.dofs, pk.kbc.values)
    s.set_segments(pk.node_begin)
    with pytest.raises(api.TahoeGpuError):
        s.set_preconditioner(api.PC_AMG)
    s.close()


@pytest.mark.parametrize("degree,weight", [(2, 100), (3, 150), (4, 190)])
def test_amg_knobs(degree, weight):
    """Chebyshev degree and coarse-correction weight change the iteration
    count, never the converged displacements."""
    from tahoe_amd import api
    g = load_golden("tl_n8")
    s = make_solid(g)
    s.set_preconditioner(api.PC_AMG)
    s.set_tuning(api.TUNE_AMG_DEGREE, degree)
    s.set_tuning(api.TUNE_AMG_WEIGHT, weight)
    assert s.tuning(api.TUNE_AMG_DEGREE) == degree and s.tuning(api.TUNE_AMG_WEIGHT) == weight
    us, iters = api.static_solve(s, int(g["num_steps"]), s.opts(lin_rtol=1e-12))
    for k in range(us.shape[0]):
        ref = g["step_u"][k]
        assert np.abs(us[k] - ref).max() <= TOL_U * np.abs(ref).max(), (degree, weight, k)
    assert list(iters) == g["step_iters"].tolist()
    with pytest.raises(api.TahoeGpuError):
        s.set_tuning(api.TUNE_AMG_WEIGHT, 200)
    s.close()


@pytest.mark.parametrize("name,precision,pc", [("rve_n4", 0, 0), ("tl_n8", 1, 0), ("tl_n8", 0, 2), ("q1p0_n4", 0, 0)])
def test_fused_pcg_iteration(name, precision, pc):
    """The fused PCG iteration (direction formed inside the SpMV) and the
    five-kernel iteration do the same arithmetic: bitwise the same
    displacements and iteration counts."""
    from tahoe_amd import api
    g = load_golden(name)
    out = []
    for fused in (1, 0):
        s = make_solid(g)
        if precision:
            s.set_precision(api.MIXED)
            s.initialize()
        s.set_preconditioner(pc)
        s.set_tuning(api.TUNE_PCG_FUSED, fused)
        assert s.tuning(api.TUNE_PCG_FUSED) == fused
        lin = []
        us, iters = api.static_solve(s, int(g["num_steps"]), s.opts(lin_rtol=1e-12),
                                     on_step=lambda k, st: lin.append(st.lin_iterations))
        out.append((us, iters, lin))
        s.close()
    (u1, n1, l1), (u0, n0, l0) = out
    assert n1 == n0 and l1 == l0
    assert np.array_equal(u1, u0)
    for k in range(u1.shape[0]):
        ref = g["step_u"][k]
        assert np.abs(u1[k] - ref).max() <= TOL_U * np.abs(ref).max()


def test_newton_stepwise_equals_newton():
    """The host-driven Newton loop (one tg_newton_correction per call, the
    norm from tg_form_rhs) makes the same corrections as tg_newton: bitwise
    equal displacements, the same correction counts per step."""
    from tahoe_amd import api
    g = load_golden("tl_n8")
    out = []
    for stepwise in (False, True):
        s = make_solid(g)
        seen = []
        us, iters, subs = api.static_solve_adaptive(s, int(g["num_steps"]), s.opts(lin_rtol=1e-12),
                                                    log=seen.append, log_corrections=stepwise)
        out.append((us, iters, subs))
        assert any("correction" in m for m in seen) == stepwise
        s.close()
    (u0, n0, s0), (u1, n1, s1) = out
    assert n0 == n1 and s0 == s1
    assert np.array_equal(u0, u1)
    for k in range(u1.shape[0]):
        ref = g["step_u"][k]
        assert np.abs(u1[k] - ref).max() <= TOL_U * np.abs(ref).max()


def test_every_dof_prescribed_and_isolated_node():
    """Edge cases of the equation system: with every dof prescribed (no
    equations) Newton is converged at once and u is the prescribed field;
    a node that no element uses keeps free dofs with an empty row (zero
    diagonal, skipped by the Jacobi preconditioner as DiagonalMatrixT skips
    small pivots) and does not disturb the solve of the others."""
    from tahoe_amd import api, mesh as M
    m = M.cube(2, seed=3)
    H = np.array([[0.01, 0.002, 0.0], [0.0, -0.005, 0.001], [0.003, 0.0, 0.008]])
    nodes = np.repeat(np.arange(m.nnd, dtype=np.int32), 3)
    dofs = np.tile(np.arange(3, dtype=np.int32), m.nnd)
    vals = (m.X @ H.T).ravel()
    s = api.Solid(0)
    s.set_mesh(m.X, m.conn, m.mat)
    s.set_material(0, *api.mu_kappa(1000.0, 0.3))
    s.set_kbc(nodes, dofs, vals)
    n_eq, _ = s.initialize()
    assert n_eq == 0
    s.apply_kbc(1.0)
    st = s.newton(s.opts())
    assert st.status == 0 and st.iterations == 0
    assert np.allclose(s.displacement(), m.X @ H.T, rtol=0, atol=1e-15)
    s.close()
    # the golden 4^3 problem plus one node that no element uses
    g = load_golden("tl_n4")
    X = np.vstack([g["X"], [[5.0, 5.0, 5.0]]])
    s = api.Solid(0)
    s.set_mesh(X, g["conn"], g["mat"])
    for k, (E, nu) in enumerate(zip(g["E"], g["nu"])):
        s.set_material(k, *api.mu_kappa(E, nu))
    s.set_kbc(g["kbc_nodes"], g["kbc_dofs"], g["kbc_values"])
    n_eq, _ = s.initialize()
    assert n_eq == len(g["eqnos"][g["eqnos"] > 0]) + 3
    us, iters = api.static_solve(s, int(g["num_steps"]), s.opts(lin_rtol=1e-12))
    for k in range(us.shape[0]):
        ref = g["step_u"][k]
        assert np.abs(us[k][:-1] - ref).max() <= TOL_U * np.abs(ref).max()
        assert np.all(us[k][-1] == 0.0)
    s.close()


def test_input_errors_are_reported():
    """Host-side checks turn malformed input into Tahoe status codes before
    any kernel runs (6 kBadInputValue, 4 kOutOfRange, 1 kGeneralFail), and a
    rejected call leaves the handle usable."""
    from tahoe_amd import api, mesh as M
    m = M.cube(2, seed=1)
    kbc = M.kbc_uniaxial(m, 0.05)
    s = api.Solid(0)

    def code(fn, *a):
        with pytest.raises(api.TahoeGpuError) as e:
            fn(*a)
        return e.value.code

    assert code(s.set_mesh, np.zeros((0, 3)), np.zeros((0, 8), dtype=np.int32)) == 6  # empty mesh
    bad = m.conn.copy()
    bad[3, 5] = m.nnd  # a node that does not exist
    assert code(s.set_mesh, m.X, bad) == 4
    rep = m.conn.copy()
    rep[0, 1] = rep[0, 0]  # repeated node
    assert code(s.set_mesh, m.X, rep) == 6
    assert code(s.set_mesh, m.X, m.conn, np.full(m.nel, -1)) == 6
    assert code(s.initialize) == 1  # no mesh yet
    s.set_mesh(m.X, m.conn, m.mat)
    assert code(s.set_material, 2000, 1.0, 1.0) == 4
    assert code(s.set_material, 0, -1.0, 1.0) == 1
    assert code(s.set_formulation, 7) == 6
    assert code(s.set_material_model, 9) == 6
    assert code(s.set_kbc, [m.nnd], [0], [0.0]) == 4
    assert code(s.set_kbc, [0], [3], [0.0]) == 4
    assert code(s.set_forces, [0, m.nnd], [0, 0], [1.0, 1.0]) == 4
    with pytest.raises(ValueError):
        s.set_kbc([0, 1], [0], [0.0])
    assert code(s.initialize) == 6  # material 0 not defined
    s.set_material(0, *api.mu_kappa(1000.0, 0.3))
    s.set_kbc(kbc.nodes, kbc.dofs, kbc.values)
    s.initialize()
    us, iters = api.static_solve(s, 1, s.opts())
    s.close()
    # the rejected calls left nothing behind: a fresh problem solves the same
    f = api.Solid(0)
    f.set_mesh(m.X, m.conn, m.mat)
    f.set_material(0, *api.mu_kappa(1000.0, 0.3))
    f.set_kbc(kbc.nodes, kbc.dofs, kbc.values)
    f.initialize()
    uf, itf = api.static_solve(f, 1, f.opts())
    f.close()
    assert iters[0] > 0 and list(iters) == list(itf) and np.array_equal(us, uf)


def test_kbc_rechecked_against_a_later_mesh():
    """Kinematic cards set on one mesh and a smaller mesh set afterwards:
    tg_initialize rejects the cards that name missing nodes (kOutOfRange)
    instead of numbering equations past the end of the node table."""
    from tahoe_amd import api, mesh as M
    big, small = M.cube(3, seed=1), M.cube(2, seed=1)
    kbc = M.kbc_uniaxial(big, 0.05)
    assert kbc.nodes.max() >= small.nnd
    s = api.Solid(0)
    s.set_mesh(big.X, big.conn, big.mat)
    s.set_material(0, *api.mu_kappa(1000.0, 0.3))
    s.set_kbc(kbc.nodes, kbc.dofs, kbc.values)
    s.set_mesh(small.X, small.conn, small.mat)
    with pytest.raises(api.TahoeGpuError) as e:
        s.initialize()
    assert e.value.code == 4
    ks = M.kbc_uniaxial(small, 0.05)
    s.set_kbc(ks.nodes, ks.dofs, ks.values)
    s.initialize()
    s.close()


def test_reinitialized_context_solves_like_a_fresh_one():
    """A context solved once (its PCG chunk captured as a graph) and then
    re-initialized on another mesh with the same node count but other
    segments solves exactly as a fresh context does: the captured chunk is
    rebuilt for the new buffers and sizes, not replayed."""
    from tahoe_amd import api, mesh as M
    m1, m2 = M.cube(4, seed=3), M.cube(4, seed=4)
    kbc = M.kbc_uniaxial(m1, 0.05)

    def define(s, m, segs=None):
        s.set_mesh(m.X, m.conn, m.mat)
        s.set_material(0, *api.mu_kappa(1000.0, 0.3))
        s.set_kbc(kbc.nodes, kbc.dofs, kbc.values)
        s.initialize()

    s = api.Solid(0)
    define(s, m1)
    api.static_solve(s, 1, s.opts())
    define(s, m2)
    u_re, it_re = api.static_solve(s, 2, s.opts())
    s.close()
    f = api.Solid(0)
    define(f, m2)
    u_f, it_f = api.static_solve(f, 2, f.opts())
    f.close()
    assert list(it_re) == list(it_f) and np.array_equal(u_re, u_f)


def test_form_tangent_assembles_newtons_tangent():
    """tg_form_tangent forms K(u) for the inner solve: through the MSR
    (assembly 0), whose values then equal the reference's K at that iterate,
    or straight into the block operator (assembly 1); either way it launches
    the stiffness pass and the operator/diagonal pass and leaves a following
    correction's result unchanged."""
    from tahoe_amd import api
    g = load_golden("tl_n4")
    n = None
    for assembly in (0, 1):
        s = make_solid(g)
        n = s.n_eq
        s.set_tuning(api.TUNE_NEWTON_ASSEMBLY, assembly)
        s.set_displacement(g["it_u"][1])
        s.timing_enable(True)
        s.timing_reset()
        s.form_tangent()
        t = s.timing()
        s.timing_enable(False)
        assert t["stiffness_launches"] > 0 and t["expand_launches"] > 0
        if assembly == 0:
            A = msr_dense(s.msr_bindx(), s.msr_values(), n)
            K = golden_K_dense(g, 1, n)
            assert np.abs(A - K).max() <= 1e-12 * np.abs(K).max()
        st = s.newton_correction(s.opts(lin_rtol=1e-12))
        u1 = s.displacement()
        s.close()
        f = make_solid(g)
        f.set_tuning(api.TUNE_NEWTON_ASSEMBLY, assembly)
        f.set_displacement(g["it_u"][1])
        st_f = f.newton_correction(f.opts(lin_rtol=1e-12))
        assert st.lin_iterations == st_f.lin_iterations and np.array_equal(u1, f.displacement())
        f.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("shape", [0, 9, 16, 17, 18])
@pytest.mark.parametrize("chunk", [0, 100])
def test_gather_assembly_equals_scatter(precision, shape, chunk):
    """The element-buffer assembly (TUNE_STIFF_GATHER = 1: element entries
    written in launch order, one gather pass per MSR value) gives the MSR
    values of the colour-ordered scatter bitwise, in FP64 and FP32, for the
    block and the matrix-core kernels and for split launch ranges."""
    from tahoe_amd import api
    g = load_golden("tl_n8")
    out = []
    for gather in (0, 1):
        s = make_solid(g) if precision == 0 else None
        if s is None:
            s = api.Solid(0)
            s.set_mesh(g["X"], g["conn"], g["mat"])
            s.set_material(0, *api.mu_kappa(1000.0, 0.3))
            s.set_precision(api.MIXED)
            s.set_kbc(g["kbc_nodes"], g["kbc_dofs"], g["kbc_values"])
            s.initialize()
        s.set_tuning(api.TUNE_STIFF_SHAPE, shape)
        s.set_tuning(api.TUNE_STIFF_CHUNK, chunk)
        s.set_tuning(api.TUNE_STIFF_GATHER, gather)
        assert s.tuning(api.TUNE_STIFF_GATHER) == gather
        s.set_displacement(g["step_u"][1])
        s.form_lhs()
        s.form_lhs()  # a second pass overwrites every value
        out.append(s.msr_values())
        s.close()
    assert np.array_equal(out[0], out[1])


def test_gather_assembly_newton_and_batch():
    """Newton with the gather assembly reaches the reference's converged
    displacements with its Newton counts; a batch of RVEs assembles the
    same values with either assembly."""
    from tahoe_amd import api, batch, rve
    g = load_golden("tl_n16")
    s = make_solid(g)
    s.set_tuning(api.TUNE_STIFF_GATHER, 1)
    s.set_tuning(api.TUNE_STIFF_SHAPE, 16)
    step_u, iters = api.static_solve(s, int(g["num_steps"]), s.opts(lin_rtol=1e-12))
    s.close()
    for k in range(step_u.shape[0]):
        ref = g["step_u"][k]
        assert np.abs(step_u[k] - ref).max() <= 1e-9 * np.abs(ref).max(), k
    assert list(iters) == g["step_iters"].tolist()
    reqs = rve.batch_requests(12, seed=77)
    pk = batch.pack([r.problem() for r in reqs])
    vals = []
    for gather in (0, 1):
        b = batch.make_solid(pk, formulation=api.TOTAL_LAGRANGIAN, precision=api.FP64)
        b.set_tuning(api.TUNE_STIFF_GATHER, gather)
        b.apply_kbc(1.0)
        b.form_lhs()
        vals.append(b.msr_values())
        b.close()
    assert np.array_equal(vals[0], vals[1])

