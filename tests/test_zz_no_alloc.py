"""tg_solve_noalloc (the solve that allocates nothing): its result equals
tg_solve's on the same matrix and right-hand side, on two calls in a row
that share the GPU's scratch array, the second on a smaller problem; sizes
beyond the scratch and the multigrid preconditioner are rejected.

Device buffers for b and x are the caller's: the test allocates them with
the HIP runtime the library itself is linked against."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    from tahoe_amd import api
    api.lib()  # loads libamdhip64 with the library
    return C.CDLL("libamdhip64.so.7")


class DevArray:
    def __init__(self, n):
        self.n = n
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), C.c_size_t(8 * n)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        assert hip().hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(8 * self.n), H2D) == 0

    def get(self):
        a = np.empty(self.n)
        assert hip().hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(8 * self.n), D2H) == 0
        return a

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()


def cube_system(n, seed, precond=0):
    from tahoe_amd import api, mesh as M
    m = M.cube(n, seed=seed)
    kbc = M.kbc_uniaxial(m, 0.05)
    s = api.Solid(0)
    s.set_mesh(m.X, m.conn, m.mat)
    s.set_material(0, *api.mu_kappa(1000.0, 0.3))
    s.set_kbc(kbc.nodes, kbc.dofs, kbc.values)
    s.set_preconditioner(precond)
    s.initialize()
    s.apply_kbc(0.2)
    s.form_lhs()
    R, _ = s.form_rhs()
    return s, R


@pytest.mark.parametrize("precond", [0, 1])
def test_noalloc_solve_equals_solve_two_calls_second_smaller(precond):
    for n, seed in ((8, 11), (4, 12)):
        s, R = cube_system(n, seed, precond)
        x_ref, st_ref = s.solve(R, rtol=1e-10, maxit=100000)
        b, x = DevArray(s.n_eq), DevArray(s.n_eq)
        try:
            b.put(R)
            x.put(np.full(s.n_eq, np.nan))   # every entry is written
            st = s.solve_noalloc(b.p.value, x.p.value, rtol=1e-10, maxit=100000)
            xn = x.get()
        finally:
            b.free()
            x.free()
            s.close()
        assert st.converged == 1 and st.iterations == st_ref.iterations, (n, st.iterations, st_ref.iterations)
        assert np.array_equal(xn, x_ref), (n, np.abs(xn - x_ref).max())


def test_noalloc_rejects_oversize_and_multigrid():
    from tahoe_amd import api
    s, R = cube_system(80, 3)   # 81^3 nodes need ~76 MB of scratch
    b, x = DevArray(s.n_eq), DevArray(s.n_eq)
    try:
        with pytest.raises(api.TahoeGpuError) as e:
            s.solve_noalloc(b.p.value, x.p.value)
        assert e.value.code == 5
    finally:
        b.free()
        x.free()
        s.close()
    s, R = cube_system(4, 5, precond=2)
    b, x = DevArray(s.n_eq), DevArray(s.n_eq)
    try:
        with pytest.raises(api.TahoeGpuError) as e:
            s.solve_noalloc(b.p.value, x.p.value)
        assert e.value.code == 6
    finally:
        b.free()
        x.free()
        s.close()
