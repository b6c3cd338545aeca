import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def golden_formulation(g):
    return 1 if "Q1P0" in str(g["element"]) else 0


def golden_R_eq(g, k):
    """Residual of golden iterate k in equation order (from nodal f)."""
    eq = g["eqnos"]
    f = g["it_f"][k]
    n = int(eq.max())
    R = np.zeros(n)
    m = eq > 0
    R[eq[m] - 1] = f[m]
    return R


def golden_K_dense(g, k, n):
    s, e = g["lhs_ptr"][k], g["lhs_ptr"][k + 1]
    A = np.zeros((n, n))
    A[g["lhs_r"][s:e] - 1, g["lhs_c"][s:e] - 1] = g["lhs_v"][s:e]
    return A


def msr_dense(bindx, val, n):
    A = np.zeros((n, n))
    A[np.arange(n), np.arange(n)] = val[:n]
    for r in range(n):
        cols = bindx[bindx[r]:bindx[r + 1]]
        A[r, cols] = val[bindx[r]:bindx[r + 1]]
    return A


def msr_pattern_dense(bindx, n):
    P = np.zeros((n, n), dtype=bool)
    P[np.arange(n), np.arange(n)] = True
    for r in range(n):
        P[r, bindx[bindx[r]:bindx[r + 1]]] = True
    return P
