"""ctypes front end of the C oracle -- TEST INFRASTRUCTURE.

Loaded only by tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
bench.py, always as the checker.  Build with ``make -C oracle``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)

TL, Q1P0, TL_STVENANT, Q1P0_STVENANT = 0, 1, 2, 3


class NewtonOpts(C.Structure):
    _fields_ = [("rel_tol", C.c_double), ("abs_tol", C.c_double), ("div_tol", C.c_double),
                ("max_iterations", C.c_int), ("lin_rtol", C.c_double), ("lin_maxit", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: run `make -C {HERE}`")
        L = C.CDLL(LIB_PATH)
        L.or_mu_kappa.argtypes = [C.c_double, C.c_double, D, D]
        L.or_simo_cauchy.argtypes = [C.c_double, C.c_double, D, D]
        L.or_simo_moduli.argtypes = [C.c_double, C.c_double, D, D]
        L.or_simo_energy.argtypes = [C.c_double, C.c_double, D]
        L.or_simo_energy.restype = C.c_double
        L.or_stvenant_cauchy.argtypes = [C.c_double, C.c_double, D, D]
        L.or_stvenant_moduli.argtypes = [C.c_double, C.c_double, D, D]
        L.or_stvenant_energy.argtypes = [C.c_double, C.c_double, D]
        L.or_stvenant_energy.restype = C.c_double
        L.or_tl_stvenant_element.argtypes = [D, D, C.c_double, C.c_double, D, D]
        L.or_hex8_parent.argtypes = [D, D, D]
        L.or_tl_element.argtypes = [D, D, C.c_double, C.c_double, D, D]
        L.or_q1p0_element.argtypes = [D, D, C.c_double, C.c_double, D, D]
        L.or_q1p0_stvenant_element.argtypes = [D, D, C.c_double, C.c_double, D, D]
        L.or_equation_numbers.argtypes = [C.c_int, C.c_int, I, I, I]
        L.or_msr_pattern.argtypes = [C.c_int, C.c_int, C.c_int, I, I, I, C.c_long]
        L.or_msr_pattern.restype = C.c_long
        L.or_residual.argtypes = [C.c_int, C.c_int, C.c_int, D, I, I, D, D, D, I, C.c_int, D]
        L.or_stiffness.argtypes = [C.c_int, C.c_int, C.c_int, D, I, I, D, D, D, I, C.c_int, I, D]
        L.or_color_elements.argtypes = [C.c_int, C.c_int, I, I, I]
        L.or_element_passes_parallel.argtypes = [C.c_int, C.c_int, D, I, I, D, D, D, I, C.c_int, I, D, D, C.c_int,
                                                 I, I, C.c_int]
        L.or_msr_multx.argtypes = [C.c_int, I, D, D, D]
        L.or_gmres_dense.argtypes = [C.c_int, D, D, D, C.c_int, C.c_double, C.c_int, I]
        L.or_gmres_msr.argtypes = [C.c_int, I, D, D, D, C.c_int, C.c_double, C.c_int, I]
        L.or_pcg_jacobi.argtypes = [C.c_int, I, D, D, D, C.c_double, C.c_int, I, D]
        L.or_static_solve.argtypes = [C.c_int, C.c_int, C.c_int, D, I, I, D, D, C.c_int, I, I, D,
                                      C.c_int, I, I, D, C.c_int, C.POINTER(NewtonOpts), D, D, I, I]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(D)


def _i(a):
    return a.ctypes.data_as(I)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def mu_kappa(E, nu):
    m, k = C.c_double(), C.c_double()
    lib().or_mu_kappa(E, nu, C.byref(m), C.byref(k))
    return m.value, k.value


def simo_cauchy(mu, kappa, F):
    """F is a 3x3 array (row, col); returns Voigt stress (11,22,33,23,13,12)."""
    Fc = f64(np.asarray(F).T.ravel())  # column-major
    s = np.zeros(6)
    rc = lib().or_simo_cauchy(mu, kappa, _d(Fc), _d(s))
    if rc:
        raise ArithmeticError(rc)
    return s


def simo_moduli(mu, kappa, F):
    Fc = f64(np.asarray(F).T.ravel())
    c = np.zeros(36)
    rc = lib().or_simo_moduli(mu, kappa, _d(Fc), _d(c))
    if rc:
        raise ArithmeticError(rc)
    return c.reshape(6, 6).T.copy()  # stored column-major


def simo_energy(mu, kappa, F):
    Fc = f64(np.asarray(F).T.ravel())
    return lib().or_simo_energy(mu, kappa, _d(Fc))


def stvenant_cauchy(mu, kappa, F):
    """large_strain_StVenant Cauchy stress, Voigt (11,22,33,23,13,12)."""
    Fc = f64(np.asarray(F).T.ravel())
    s = np.zeros(6)
    rc = lib().or_stvenant_cauchy(mu, kappa, _d(Fc), _d(s))
    if rc:
        raise ArithmeticError(rc)
    return s


def stvenant_moduli(mu, kappa, F):
    Fc = f64(np.asarray(F).T.ravel())
    c = np.zeros(36)
    rc = lib().or_stvenant_moduli(mu, kappa, _d(Fc), _d(c))
    if rc:
        raise ArithmeticError(rc)
    return c.reshape(6, 6).T.copy()


def stvenant_energy(mu, kappa, F):
    Fc = f64(np.asarray(F).T.ravel())
    return lib().or_stvenant_energy(mu, kappa, _d(Fc))


def hex8_parent():
    Na = np.zeros((8, 8)); DNa = np.zeros((8, 3, 8)); w = np.zeros(8)
    lib().or_hex8_parent(_d(Na), _d(DNa), _d(w))
    return Na, DNa, w


def element(kind, Xe, ue, mu, kappa, want_R=True, want_K=True):
    Xe = f64(Xe).ravel(); ue = f64(ue).ravel()
    R = np.zeros(24) if want_R else None
    K = np.zeros((24, 24)) if want_K else None
    fn = {Q1P0: lib().or_q1p0_element, TL_STVENANT: lib().or_tl_stvenant_element,
          Q1P0_STVENANT: lib().or_q1p0_stvenant_element}.get(kind, lib().or_tl_element)
    rc = fn(_d(Xe), _d(ue), mu, kappa, _d(R) if want_R else None, _d(K) if want_K else None)
    if rc:
        raise ArithmeticError(rc)
    return R, K


def equation_numbers(nnd, kbc_nodes, kbc_dofs):
    eq = np.zeros((nnd, 3), dtype=np.int32)
    kn, kd = i32(kbc_nodes), i32(kbc_dofs)
    n = lib().or_equation_numbers(nnd, len(kn), _i(kn), _i(kd), _i(eq))
    return eq, n


def msr_pattern(n_eq, conn, eqnos):
    conn = i32(conn); eqnos = i32(eqnos)
    nel = conn.shape[0]; nnd = eqnos.shape[0]
    L = lib().or_msr_pattern(n_eq, nnd, nel, _i(conn), _i(eqnos), None, 0)
    b = np.zeros(L, dtype=np.int32)
    lib().or_msr_pattern(n_eq, nnd, nel, _i(conn), _i(eqnos), _i(b), L)
    return b


def _mats(mu, kappa):
    return f64(np.atleast_1d(mu)), f64(np.atleast_1d(kappa))


def residual(kind, X, conn, mat, mu, kappa, u, eqnos, n_eq):
    X, conn, mat, u, eqnos = f64(X), i32(conn), i32(mat), f64(u), i32(eqnos)
    mu, kappa = _mats(mu, kappa)
    R = np.zeros(n_eq)
    rc = lib().or_residual(kind, X.shape[0], conn.shape[0], _d(X), _i(conn), _i(mat), _d(mu), _d(kappa),
                           _d(u), _i(eqnos), n_eq, _d(R))
    if rc:
        raise ArithmeticError(rc)
    return R


def stiffness(kind, X, conn, mat, mu, kappa, u, eqnos, n_eq, bindx):
    X, conn, mat, u, eqnos, bindx = f64(X), i32(conn), i32(mat), f64(u), i32(eqnos), i32(bindx)
    mu, kappa = _mats(mu, kappa)
    val = np.zeros(len(bindx))
    rc = lib().or_stiffness(kind, X.shape[0], conn.shape[0], _d(X), _i(conn), _i(mat), _d(mu), _d(kappa),
                            _d(u), _i(eqnos), n_eq, _i(bindx), _d(val))
    if rc:
        raise ArithmeticError(rc)
    return val


def color_elements(nnd, conn):
    """Greedy element colouring: (colour pointer, elements by colour)."""
    conn = i32(conn)
    ptr = np.zeros(65, dtype=np.int32)
    elems = np.zeros(conn.shape[0], dtype=np.int32)
    nc = lib().or_color_elements(nnd, conn.shape[0], _i(conn), _i(ptr), _i(elems))
    if nc < 0:
        raise RuntimeError("colouring failed")
    return ptr[:nc + 1].copy(), elems


def element_passes_parallel(kind, X, conn, mat, mu, kappa, u, eqnos, n_eq, bindx, colors, nthreads,
                            want_R=True, want_K=True):
    """Residual and stiffness passes over the colouring on nthreads OpenMP
    threads (the multi-threaded CPU baseline); returns (R, val)."""
    X, conn, mat, u, eqnos, bindx = f64(X), i32(conn), i32(mat), f64(u), i32(eqnos), i32(bindx)
    mu, kappa = _mats(mu, kappa)
    ptr, elems = colors
    R = np.zeros(n_eq) if want_R else None
    val = np.zeros(len(bindx)) if want_K else None
    rc = lib().or_element_passes_parallel(kind, conn.shape[0], _d(X), _i(conn), _i(mat), _d(mu), _d(kappa), _d(u),
                                          _i(eqnos), n_eq, _i(bindx), _d(R) if want_R else None,
                                          _d(val) if want_K else None, len(ptr) - 1, _i(ptr), _i(elems), nthreads)
    if rc:
        raise ArithmeticError(rc)
    return R, val


def msr_multx(bindx, val, x):
    n = int(bindx[0]) - 1
    bindx, val, x = i32(bindx), f64(val), f64(x)
    b = np.zeros(n)
    lib().or_msr_multx(n, _i(bindx), _d(val), _d(x), _d(b))
    return b


def pcg_jacobi(bindx, val, b, rtol=1e-12, maxit=100000):
    n = int(bindx[0]) - 1
    bindx, val, b = i32(bindx), f64(val), f64(b)
    x = np.zeros(n)
    its, rr = C.c_int(), C.c_double()
    rc = lib().or_pcg_jacobi(n, _i(bindx), _d(val), _d(b), _d(x), rtol, maxit, C.byref(its), C.byref(rr))
    return x, its.value, rr.value, rc


def static_solve(kind, X, conn, mat, mu, kappa, kbc_nodes, kbc_dofs, kbc_values, num_steps,
                 rel_tol=1e-10, abs_tol=1e-14, div_tol=10.0, max_iterations=20, lin_rtol=1e-12,
                 lin_maxit=200000, force_nodes=(), force_dofs=(), force_values=()):
    """FEManagerT::Solve's static load path: num_steps steps of t = s /
    num_steps, kinematic cards at value * t, force cards adding value * t to
    the residual, Newton per step.  Returns status, final u, u and Newton
    count per step, total PCG iterations."""
    X, conn, mat = f64(X), i32(conn), i32(mat)
    kn, kd, kv = i32(kbc_nodes), i32(kbc_dofs), f64(kbc_values)
    fn, fd, fv = i32(force_nodes), i32(force_dofs), f64(force_values)
    if not (len(fn) == len(fd) == len(fv)):
        raise ValueError("force card arrays differ in length")
    mu, kappa = _mats(mu, kappa)
    nnd = X.shape[0]
    u = np.zeros((nnd, 3))
    step_u = np.zeros((num_steps, nnd, 3))
    iters = np.zeros(num_steps, dtype=np.int32)
    lin = C.c_int()
    o = NewtonOpts(rel_tol, abs_tol, div_tol, max_iterations, lin_rtol, lin_maxit)
    rc = lib().or_static_solve(kind, nnd, conn.shape[0], _d(X), _i(conn), _i(mat), _d(mu), _d(kappa),
                               len(kn), _i(kn), _i(kd), _d(kv), len(fn), _i(fn), _i(fd), _d(fv), num_steps,
                               C.byref(o), _d(u), _d(step_u),
                               _i(iters), C.byref(lin))
    return dict(status=rc, u=u, step_u=step_u, step_iters=iters, lin_iters=lin.value)


def gmres_dense(A, b, m, tol, maxit=0, x0=None):
    """NLSolver_NK::GMRES on a dense matrix: returns (x, converged, Arnoldi steps)."""
    A = f64(A)
    n = A.shape[0]
    b = f64(b)
    x = f64(np.zeros(n) if x0 is None else x0).copy()
    st = C.c_int()
    rc = lib().or_gmres_dense(n, _d(A), _d(b), _d(x), m, tol, maxit, C.byref(st))
    return x, bool(rc), st.value


def gmres_msr(bindx, val, b, m, tol, maxit=0):
    n = int(bindx[0]) - 1
    bindx, val, b = i32(bindx), f64(val), f64(b)
    x = np.zeros(n)
    st = C.c_int()
    rc = lib().or_gmres_msr(n, _i(bindx), _d(val), _d(b), _d(x), m, tol, maxit, C.byref(st))
    return x, bool(rc), st.value
