"""Mixed precision (FP32 stiffness and inner Jacobi-PCG, FP64 residual,
Newton test and update): the assembled FP32 stiffness is within FP32
rounding of the FP64 one, and Newton driven by the FP64 residual reaches
the FP64 solution -- converged displacements within 1e-9 relative of the
reference fixtures, as the FP64 path."""
import numpy as np
import pytest

from conftest import load_golden, msr_dense
from test_gpu_tl import make_solid

pytestmark = pytest.mark.gpu


def mixed_solid(g):
    from tahoe_amd import api
    s = api.Solid(0)
    s.set_mesh(g["X"], g["conn"], g["mat"])
    for m, (E, nu) in enumerate(zip(g["E"], g["nu"])):
        s.set_material(m, *api.mu_kappa(E, nu))
    s.set_formulation(api.TOTAL_LAGRANGIAN)
    s.set_precision(api.MIXED)
    s.set_kbc(g["kbc_nodes"], g["kbc_dofs"], g["kbc_values"])
    s.initialize()
    return s


@pytest.mark.parametrize("shape", [0, 6, 9, 15, 16, 17, 18])
def test_mixed_stiffness_is_fp32_rounding_of_fp64(shape):
    """In several stiffness shapes: element-major lanes (0), the D*B form
    (6), loads issued before the block products (9), block-major lanes with
    atomic additions (15), FP32 matrix-core tiles with 8 and 4 elements per
    wave (16, 17)."""
    from tahoe_amd import api
    g = load_golden("tl_n8")
    s64, s32 = make_solid(g), mixed_solid(g)
    s32.set_tuning(api.TUNE_STIFF_SHAPE, shape)
    u = g["step_u"][2]
    for s in (s64, s32):
        s.set_displacement(u)
        s.form_lhs()
    v64, v32 = s64.msr_values(), s32.msr_values()
    scale = np.abs(v64).max()
    err = np.abs(v32 - v64).max() / scale
    assert 0.0 < err <= 2e-6, err
    # the residual is FP64 in both modes
    R64, _ = s64.form_rhs()
    R32, _ = s32.form_rhs()
    assert np.array_equal(R64, R32)


@pytest.mark.parametrize("shape", [9, 16])
@pytest.mark.parametrize("name", ["tl_n8", "tl_n4_s20", "rve_n4", "tl_n16"])
def test_mixed_newton_reaches_fp64_solution(name, shape):
    """Newton on the FP64 residual with the FP32 tangent of the block
    kernel (shape 9) and of the matrix-core kernel (shape 16)."""
    from tahoe_amd import api
    g = load_golden(name)
    s = mixed_solid(g)
    s.set_tuning(api.TUNE_STIFF_SHAPE, shape)
    opts = api.Solid.opts(lin_rtol=1e-6)
    step_u, iters = api.static_solve(s, int(g["num_steps"]), opts)
    for k in range(step_u.shape[0]):
        ref = g["step_u"][k]
        err = np.abs(step_u[k] - ref).max() / np.abs(ref).max()
        assert err <= 1e-9, (name, k, err)
