"""ctypes binding of libtahoe_gpu.so (C ABI in include/tahoe_gpu.h).

The library must have been built in-tree (``python -m tahoe_amd.build``);
loading fails loudly otherwise -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtahoe_gpu.so")

TOTAL_LAGRANGIAN, Q1P0 = 0, 1
MAT_SIMO_ISOTROPIC, MAT_STVENANT = 0, 1
FP64, MIXED = 0, 1
TUNE_STIFF_CHUNK, TUNE_STIFF_SHAPE, TUNE_NEWTON_ASSEMBLY, TUNE_PCG_GRAPH, TUNE_SPMV_XCD = 1, 2, 3, 4, 5
TUNE_AMG_DEGREE, TUNE_AMG_WEIGHT, TUNE_PCG_FUSED, TUNE_STIFF_GATHER, TUNE_PCG_PERSIST = 6, 7, 8, 9, 10
TUNE_AMG_REUSE = 11
PC_JACOBI, PC_BLOCK_JACOBI, PC_AMG = 0, 1, 2
LIN_PCG, LIN_GMRES = 0, 1
STIFF_SHAPE_EPB = (16, 12, 8, 4, 16, 8, 16, 8, 4, 12, 8, 4, 12, 8, 4, 16, 8, 4, 8)
STIFF_SHAPE_BLOCK_MAJOR = (False,) * 4 + (True, True) + (False,) * 9 + (True,) + (False,) * 3
STIFF_SHAPE_DB_FORM = (False,) * 6 + (True,) * 3 + (False,) * 10
# 0 load-add-store, 1 loads first, 2 atomic adds, 3 matrix-core tiles (MSR path; elements per wave),
# 4 matrix-core tiles with registers bounded for 3 waves per SIMD
STIFF_SHAPE_SCATTER = (0,) * 9 + (1,) * 3 + (2,) * 4 + (3,) * 2 + (4,)
STIFF_SHAPE_MFMA = tuple(x >= 3 for x in STIFF_SHAPE_SCATTER)

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)

# exported symbols, in declaration order of include/tahoe_gpu.h
SYMBOLS = [
    "tg_create", "tg_destroy", "tg_last_error", "tg_set_mesh", "tg_set_material", "tg_set_formulation",
    "tg_set_material_model", "tg_set_precision", "tg_set_preconditioner", "tg_set_linear_solver", "tg_set_kbc", "tg_set_forces",
    "tg_apply_forces", "tg_set_segments", "tg_initialize", "tg_get_eqnos", "tg_get_msr_bindx", "tg_set_displacement",
    "tg_get_displacement", "tg_apply_kbc", "tg_close_step", "tg_reset_step", "tg_form_rhs", "tg_form_lhs",
    "tg_form_tangent", "tg_get_msr_values", "tg_multx", "tg_copy_diagonal", "tg_solve", "tg_solve_noalloc", "tg_newton", "tg_newton_batch", "tg_newton_correction",
    "tg_synchronize", "tg_timing_enable", "tg_timing_reset", "tg_timing_get", "tg_calib_copy", "tg_set_tuning",
    "tg_get_tuning",
]


class LinStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("rel_residual", C.c_double), ("converged", C.c_int)]


class NewtonOpts(C.Structure):
    _fields_ = [("rel_tolerance", C.c_double), ("abs_tolerance", C.c_double),
                ("divergence_tolerance", C.c_double), ("max_iterations", C.c_int),
                ("lin_rtol", C.c_double), ("lin_maxit", C.c_int)]


class NewtonStats(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("error0", C.c_double), ("error", C.c_double),
                ("lin_iterations", C.c_long)]


class Timing(C.Structure):
    _fields_ = [("residual_ms", C.c_double), ("residual_launches", C.c_long),
                ("stiffness_ms", C.c_double), ("stiffness_launches", C.c_long),
                ("expand_ms", C.c_double), ("expand_launches", C.c_long),
                ("spmv_ms", C.c_double), ("spmv_launches", C.c_long),
                ("vector_ms", C.c_double), ("vector_launches", C.c_long),
                ("precond_ms", C.c_double), ("precond_launches", C.c_long)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TahoeGpuError(RuntimeError):
    """Non-zero status from the library; ``code`` is the ExceptionT::CodeT value."""

    def __init__(self, code, msg):
        super().__init__(f"status {code}: {msg}")
        self.code = code


_lib = None
_lib_lock = threading.Lock()


def lib():
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m tahoe_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.tg_create.argtypes = [C.POINTER(vp), C.c_int]
        L.tg_destroy.argtypes = [vp]
        L.tg_last_error.argtypes = [vp]
        L.tg_last_error.restype = C.c_char_p
        L.tg_set_mesh.argtypes = [vp, C.c_int, C.c_int, D, I, I]
        L.tg_set_material.argtypes = [vp, C.c_int, C.c_double, C.c_double]
        L.tg_set_formulation.argtypes = [vp, C.c_int]
        L.tg_set_material_model.argtypes = [vp, C.c_int]
        L.tg_set_forces.argtypes = [vp, C.c_int, I, I, D]
        L.tg_apply_forces.argtypes = [vp, C.c_double]
        L.tg_set_precision.argtypes = [vp, C.c_int]
        L.tg_set_preconditioner.argtypes = [vp, C.c_int]
        L.tg_set_linear_solver.argtypes = [vp, C.c_int, C.c_int]
        L.tg_set_kbc.argtypes = [vp, C.c_int, I, I, D]
        L.tg_set_segments.argtypes = [vp, C.c_int, I]
        L.tg_initialize.argtypes = [vp, I, C.POINTER(C.c_long)]
        L.tg_get_eqnos.argtypes = [vp, I]
        L.tg_get_msr_bindx.argtypes = [vp, I]
        L.tg_set_displacement.argtypes = [vp, D]
        L.tg_get_displacement.argtypes = [vp, D]
        L.tg_apply_kbc.argtypes = [vp, C.c_double]
        L.tg_close_step.argtypes = [vp]
        L.tg_reset_step.argtypes = [vp]
        L.tg_form_rhs.argtypes = [vp, D, D]
        L.tg_form_lhs.argtypes = [vp]
        L.tg_form_tangent.argtypes = [vp]
        L.tg_get_msr_values.argtypes = [vp, D]
        L.tg_multx.argtypes = [vp, D, D]
        L.tg_copy_diagonal.argtypes = [vp, D]
        L.tg_solve.argtypes = [vp, D, C.c_double, C.c_int, C.POINTER(LinStats)]
        L.tg_solve_noalloc.argtypes = [vp, vp, vp, C.c_double, C.c_int, C.POINTER(LinStats)]
        L.tg_newton.argtypes = [vp, C.POINTER(NewtonOpts), C.POINTER(NewtonStats)]
        L.tg_newton_correction.argtypes = [vp, C.POINTER(NewtonOpts), C.POINTER(NewtonStats)]
        L.tg_newton_batch.argtypes = [vp, C.POINTER(NewtonOpts), C.POINTER(NewtonStats)]
        L.tg_synchronize.argtypes = [vp]
        L.tg_timing_enable.argtypes = [vp, C.c_int]
        L.tg_timing_reset.argtypes = [vp]
        L.tg_timing_get.argtypes = [vp, C.POINTER(Timing)]
        L.tg_calib_copy.argtypes = [C.c_int, C.c_long, C.c_int, D]
        L.tg_set_tuning.argtypes = [vp, C.c_int, C.c_long]
        L.tg_get_tuning.argtypes = [vp, C.c_int, C.POINTER(C.c_long)]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(D)


def _i(a):
    return a.ctypes.data_as(I)


def calib_copy(nbytes, reps=10, device=0):
    """Mean time (ms) of a plain copy kernel over nbytes (reads nbytes, writes nbytes)."""
    ms = C.c_double()
    rc = lib().tg_calib_copy(device, int(nbytes), reps, C.byref(ms))
    if rc:
        raise TahoeGpuError(rc, "tg_calib_copy failed")
    return ms.value


def init_device(device=0):
    """Make the process's first HIP calls (runtime start-up, device
    selection, a stream and a few allocations) on the calling thread, so that
    worker threads created afterwards each start on an initialised runtime
    instead of racing to start it."""
    Solid(device).close()


class Solid:
    """One finite-element problem resident on a GPU (a tg_ctx handle)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        rc = lib().tg_create(C.byref(self._h), device)
        if rc:
            raise TahoeGpuError(rc, "tg_create failed")
        self.n_eq = None
        self.msr_len = None

    def close(self):
        if self._h:
            lib().tg_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ok(self, rc):
        if rc:
            raise TahoeGpuError(rc, lib().tg_last_error(self._h).decode())

    # ---- definition ----
    def set_mesh(self, X, conn, mat=None):
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 3)
        conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1, 8)
        m = None if mat is None else np.ascontiguousarray(mat, dtype=np.int32).ravel()
        if m is not None and m.shape[0] != conn.shape[0]:
            raise ValueError("one material id per element")
        self._ok(lib().tg_set_mesh(self._h, X.shape[0], conn.shape[0], _d(X), _i(conn),
                                   None if m is None else _i(m)))
        # kept only once the library has accepted the mesh
        self._X, self._conn = X, conn
        self.nnd, self.nel = X.shape[0], conn.shape[0]

    def set_material(self, mid, mu, kappa):
        self._ok(lib().tg_set_material(self._h, mid, float(mu), float(kappa)))

    def set_formulation(self, f):
        self._ok(lib().tg_set_formulation(self._h, f))

    def set_material_model(self, model):
        self._ok(lib().tg_set_material_model(self._h, int(model)))

    def set_linear_solver(self, solver, restart=30):
        self._ok(lib().tg_set_linear_solver(self._h, int(solver), int(restart)))

    def set_forces(self, nodes, dofs, values):
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        dofs = np.ascontiguousarray(dofs, dtype=np.int32)
        values = np.ascontiguousarray(values, dtype=np.float64)
        if not len(nodes) == len(dofs) == len(values):
            raise ValueError("nodes, dofs and values must have one entry per force card")
        self._ok(lib().tg_set_forces(self._h, len(nodes), _i(nodes), _i(dofs), _d(values)))

    def apply_forces(self, scale):
        self._ok(lib().tg_apply_forces(self._h, float(scale)))

    def set_precision(self, p):
        self._ok(lib().tg_set_precision(self._h, p))

    def set_preconditioner(self, pc):
        """PC_JACOBI (the reference's DiagonalMatrixT) or PC_BLOCK_JACOBI."""
        self._ok(lib().tg_set_preconditioner(self._h, pc))

    def set_kbc(self, nodes, dofs, values):
        n = np.ascontiguousarray(nodes, dtype=np.int32)
        d = np.ascontiguousarray(dofs, dtype=np.int32)
        v = np.ascontiguousarray(values, dtype=np.float64)
        if not len(n) == len(d) == len(v):
            raise ValueError("nodes, dofs and values must have one entry per kinematic card")
        self._ok(lib().tg_set_kbc(self._h, len(n), _i(n), _i(d), _d(v)))

    def set_segments(self, node_begin):
        nb = np.ascontiguousarray(node_begin, dtype=np.int32)
        self.nseg = len(nb) - 1
        self._ok(lib().tg_set_segments(self._h, self.nseg, _i(nb)))

    def initialize(self):
        n = C.c_int()
        L = C.c_long()
        self._ok(lib().tg_initialize(self._h, C.byref(n), C.byref(L)))
        self.n_eq, self.msr_len = n.value, L.value
        return self.n_eq, self.msr_len

    def eqnos(self):
        e = np.zeros((self.nnd, 3), dtype=np.int32)
        self._ok(lib().tg_get_eqnos(self._h, _i(e)))
        return e

    def msr_bindx(self):
        b = np.zeros(self.msr_len, dtype=np.int32)
        self._ok(lib().tg_get_msr_bindx(self._h, _i(b)))
        return b

    # ---- state ----
    def set_displacement(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self._ok(lib().tg_set_displacement(self._h, _d(u)))

    def displacement(self):
        u = np.zeros((self.nnd, 3))
        self._ok(lib().tg_get_displacement(self._h, _d(u)))
        return u

    def apply_kbc(self, t):
        self._ok(lib().tg_apply_kbc(self._h, float(t)))

    def close_step(self):
        self._ok(lib().tg_close_step(self._h))

    def reset_step(self):
        self._ok(lib().tg_reset_step(self._h))

    # ---- element group / matrix ----
    def form_rhs(self, want=True):
        R = np.zeros(self.n_eq) if want else None
        nrm = C.c_double()
        self._ok(lib().tg_form_rhs(self._h, _d(R) if want else None, C.byref(nrm)))
        return R, nrm.value

    def form_rhs_device(self):
        self._ok(lib().tg_form_rhs(self._h, None, None))

    def form_lhs(self):
        self._ok(lib().tg_form_lhs(self._h))

    def form_tangent(self):
        """K(u) into the inner solver's operator plus the preconditioner's
        data (the tangent half of a Newton correction)."""
        self._ok(lib().tg_form_tangent(self._h))

    def msr_values(self):
        v = np.zeros(self.msr_len)
        self._ok(lib().tg_get_msr_values(self._h, _d(v)))
        return v

    def multx(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        b = np.zeros(self.n_eq)
        self._ok(lib().tg_multx(self._h, _d(x), _d(b)))
        return b

    def copy_diagonal(self):
        d = np.zeros(self.n_eq)
        self._ok(lib().tg_copy_diagonal(self._h, _d(d)))
        return d

    def solve(self, b, rtol=1e-12, maxit=100000):
        x = np.array(b, dtype=np.float64, copy=True)
        st = LinStats()
        self._ok(lib().tg_solve(self._h, _d(x), rtol, maxit, C.byref(st)))
        return x, st

    def solve_noalloc(self, b_dev, x_dev, rtol=1e-12, maxit=100000):
        """tg_solve_noalloc: b_dev and x_dev are device addresses (ints) of
        [n_eq] FP64 arrays the caller owns; the library allocates nothing."""
        st = LinStats()
        self._ok(lib().tg_solve_noalloc(self._h, C.c_void_p(b_dev), C.c_void_p(x_dev), rtol, maxit, C.byref(st)))
        return st

    # ---- solver ----
    @staticmethod
    def opts(rel_tol=1e-10, abs_tol=1e-14, div_tol=10.0, max_iterations=20, lin_rtol=1e-12, lin_maxit=200000):
        return NewtonOpts(rel_tol, abs_tol, div_tol, max_iterations, lin_rtol, lin_maxit)

    def newton(self, opts=None):
        o = opts or self.opts()
        st = NewtonStats()
        rc = lib().tg_newton(self._h, C.byref(o), C.byref(st))
        self._ok(rc)
        return st

    def newton_stepwise(self, opts=None, log=None):
        """tg_newton's loop (NLSolver::Solve with ExitIteration,
        NLSolver.cpp:57-264,675-757) driven from the host one correction at
        a time (tg_newton_correction, then tg_form_rhs for ||R||), so that
        log(correction, ||R||/||R0||, PCG iterations) can report each one on
        a long solve.  Same tests and the same limits as tg_newton; a
        Jacobian J <= 0 raises TahoeGpuError (code 7)."""
        o = opts or self.opts()
        _, err0 = self.form_rhs(want=False)
        out = NewtonStats(0, 0, err0, err0, 0)
        if err0 < o.abs_tolerance:
            return out
        it, lin = -1, 0
        while True:
            st = self.newton_correction(o)
            it += 1
            lin += st.lin_iterations
            _, err = self.form_rhs(want=False)
            out.error, out.iterations, out.lin_iterations = err, it + 1, lin
            rel = err / err0
            if log:
                log(it + 1, rel, st.lin_iterations)
            if rel > o.divergence_tolerance:
                out.status = 1
                return out
            if rel < o.rel_tolerance or err < o.abs_tolerance:
                return out
            if it >= o.max_iterations:
                out.status = 1
                return out

    def newton_batch(self, opts=None):
        o = opts or self.opts()
        st = (NewtonStats * getattr(self, "nseg", 1))()
        self._ok(lib().tg_newton_batch(self._h, C.byref(o), st))
        return list(st)

    def newton_correction(self, opts=None):
        o = opts or self.opts()
        st = NewtonStats()
        self._ok(lib().tg_newton_correction(self._h, C.byref(o), C.byref(st)))
        return st

    def synchronize(self):
        self._ok(lib().tg_synchronize(self._h))

    def timing_enable(self, on=True):
        self._ok(lib().tg_timing_enable(self._h, 1 if on else 0))

    def timing_reset(self):
        self._ok(lib().tg_timing_reset(self._h))

    def timing(self):
        t = Timing()
        self._ok(lib().tg_timing_get(self._h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, knob, value):
        self._ok(lib().tg_set_tuning(self._h, knob, int(value)))

    def tuning(self, knob):
        v = C.c_long()
        self._ok(lib().tg_get_tuning(self._h, knob, C.byref(v)))
        return v.value


def static_solve(solid: Solid, num_steps: int, opts=None, on_step=None):
    """FEManagerT::Solve's static time loop (tahoe/src/main/FEManagerT.cpp:138-269):
    for t = s/num_steps apply the kinematic BCs and run Newton; returns the
    converged displacement per step and the Newton iteration counts."""
    out_u, iters = [], []
    for s in range(1, num_steps + 1):
        solid.apply_kbc(s / num_steps)
        st = solid.newton(opts)
        if st.status != 0:
            raise TahoeGpuError(1, f"Newton failed at step {s}: status {st.status}")
        solid.close_step()
        iters.append(st.iterations)
        out_u.append(solid.displacement())
        if on_step:
            on_step(s, st)
    return np.array(out_u), iters


class LoadPath:
    """FEManagerT::Solve's static time loop to the checkpoints t = k/num_steps
    with step cutting inside each interval (FEManagerT.cpp:243-266, with
    ResetStep): a sub-step whose Newton iteration fails (J <= 0, divergence,
    iteration limit) is reset and retried at half the increment, and a
    converged sub-step doubles the next one, never past the next checkpoint.
    advance() solves to the next checkpoint, so several paths (say one per
    precision) can be driven side by side.  log_corrections: Newton runs one
    correction per call (Solid.newton_stepwise) and log reports each."""

    def __init__(self, solid: Solid, num_steps: int, opts=None, min_dt=1.0 / 4096, log=None, log_corrections=False):
        self.solid, self.num_steps, self.opts = solid, num_steps, opts
        self.min_dt, self.log, self.log_corrections = min_dt, log, log_corrections
        self.t, self.dt, self.k = 0.0, 1.0 / num_steps, 0

    def advance(self):
        """Solve to the next checkpoint; returns (u, Newton corrections,
        sub-steps, PCG iterations) of that checkpoint."""
        solid, log = self.solid, self.log
        self.k += 1
        tk = self.k / self.num_steps
        n_newton, n_sub, lin = 0, 0, 0
        while self.t < tk - 1e-12:
            t, dt = self.t, min(self.dt, tk - self.t)
            self.dt = dt
            solid.apply_kbc(t + dt)
            try:
                if self.log_corrections and log:
                    st = solid.newton_stepwise(
                        self.opts, log=lambda i, rel, pcg: log(f"t={t:.5g}+{dt:.5g} correction {i}: "
                                                               f"|R|/|R0| {rel:.3e}, PCG {pcg}"))
                else:
                    st = solid.newton(self.opts)
                ok = st.status == 0
            except TahoeGpuError as e:
                if e.code != 7:
                    raise
                ok = False
            if log:
                log(f"t={t:.5g}+{dt:.5g}: {'converged' if ok else 'cut'}")
            if ok:
                solid.close_step()
                self.t += dt
                n_newton += st.iterations
                lin += st.lin_iterations
                n_sub += 1
                self.dt = dt * 2.0
            else:
                solid.reset_step()
                self.dt = dt * 0.5
                if self.dt < self.min_dt:
                    raise TahoeGpuError(1, f"load step below {self.min_dt} at t = {self.t}")
        self.t = tk
        return solid.displacement(), n_newton, n_sub, lin


def static_solve_adaptive(solid: Solid, num_steps: int, opts=None, on_step=None, min_dt=1.0 / 4096, log=None,
                          last=None, log_corrections=False):
    """LoadPath to every checkpoint (or the first `last`).  Returns the
    converged displacement at every checkpoint, the Newton corrections
    summed per checkpoint and the sub-step counts."""
    path = LoadPath(solid, num_steps, opts, min_dt, log, log_corrections)
    out_u, iters, subs = [], [], []
    for k in range(1, (num_steps if last is None else min(last, num_steps)) + 1):
        u, n_newton, n_sub, lin = path.advance()
        iters.append(n_newton)
        subs.append(n_sub)
        out_u.append(u)
        if on_step:
            on_step(k, NewtonStats(0, n_newton, 0.0, 0.0, lin))
    return np.array(out_u), iters, subs


def mu_kappa(E: float, nu: float):
    """Shear and bulk moduli from Young's modulus and Poisson's ratio
    (IsotropicT::Set_E_nu, tahoe/src/elements/continuum/solid/materials/
    primitives/IsotropicT.cpp:32-45)."""
    if E < 0.0 or nu > 0.5 or nu < -1.0:
        raise TahoeGpuError(1, "bad E/nu")  # IsotropicT.cpp:38-39
    mu = 0.5 * E / (1.0 + nu)
    lam = 2.0 * mu * nu / (1.0 - 2.0 * nu)
    return mu, lam + 2.0 / 3.0 * mu
