"""Synthetic micro-scale RVE requests (FE^2-style) for the batched and served
solves.

Every request is a unit cube of L^3 voxel hexahedra, two-phase
neo-Hookean (phase 2 with probability p2 per element, stiffness ratio
`ratio`), loaded by the affine boundary displacement u = H X on all outer
nodes with H_ij ~ U(-hmax, hmax).  The Newton predictor is the affine field
u = H X on every node (the Taylor assumption of FE^2 homogenisation).

Near-incompressible requests (Q1P0, nu = 0.499) take the isochoric part of
the macro deformation, F = det(I+H)^(-1/3) (I+H): a macro material point of a
near-incompressible solid does not change volume.  Imposing the raw F there
(volumetric strain tr H up to 0.06 against kappa/mu = 500) drives the mean
pressure to ~100 E, far above the shear modulus, where the tangent is
indefinite (loss of material stability: Jacobi-PCG needs ~16k iterations on
a 7^3 cube and the first correction inverts elements).  With the isochoric F
each such request converges quadratically in one load step.

Draws come from the counter-based SplitMix64 of tahoe_amd.mesh with one
stream per request.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import api, batch, mesh as M


@dataclass
class Request:
    rid: int
    L: int
    formulation: int   # api.TOTAL_LAGRANGIAN or api.Q1P0
    nu: float
    seed: int
    arrival: float = 0.0
    isochoric: bool = False  # impose det(I+H)^(-1/3) (I+H) instead of I+H

    @property
    def elements(self) -> int:
        return self.L ** 3

    def problem(self, E1=1000.0, ratio=10.0, p2=0.3, hmax=0.02) -> batch.Problem:
        """The request's mesh, phases, boundary values and predictor (built
        once per request and kept: a request carries its problem)."""
        key = (E1, ratio, p2, hmax)
        cached = self.__dict__.get("_problem")
        if cached is not None and cached[0] == key:
            return cached[1]
        p = self._build(E1, ratio, p2, hmax)
        self.__dict__["_problem"] = (key, p)
        return p

    def _build(self, E1, ratio, p2, hmax) -> batch.Problem:
        m = M.cube(self.L, seed=self.seed, perturb=0.0)
        ph = M.uniform01(self.seed, np.arange(m.nel, dtype=np.uint64)) < p2
        m.mat = ph.astype(np.int32)
        hu = M.uniform01(self.seed, np.uint64(1 << 40) + np.arange(9, dtype=np.uint64))
        H = ((2.0 * hu - 1.0) * hmax).reshape(3, 3)
        if self.isochoric:
            F = np.eye(3) + H
            H = F / np.cbrt(np.linalg.det(F)) - np.eye(3)
        # predictor: the affine field u = H X everywhere (exact for one phase)
        return batch.Problem(m, M.kbc_affine(m, H), [E1, ratio * E1], [self.nu, self.nu], m.X @ H.T)


def log_uniform_edges(n: int, lo: int, hi: int, seed: int) -> np.ndarray:
    """Edge lengths round(exp(U(ln lo, ln hi))), one per request."""
    u = M.uniform01(seed, np.uint64(1 << 41) + np.arange(n, dtype=np.uint64))
    return np.clip(np.rint(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))), lo, hi).astype(int)


def batch_requests(n=4096, lo=4, hi=24, seed=4242, nu=0.3):
    """4096 RVE cubes, edge log-uniform in 4-24 elements, compressible Q1."""
    L = log_uniform_edges(n, lo, hi, seed)
    return [Request(i, int(L[i]), api.TOTAL_LAGRANGIAN, nu, seed * 100003 + i) for i in range(n)]


def stream_requests(n, rate, lo=4, hi=32, seed=777, q1p0_fraction=0.5, isochoric=True):
    """Poisson arrivals (exponential gaps at `rate` per second), edge
    log-uniform in lo-hi, a fraction q1p0_fraction of Q1P0 mixed-pressure
    requests at nu = 0.499 (the isochoric part of the macro F, or the raw
    F = I + H with isochoric=False) and the rest Q1 at nu = 0.3."""
    L = log_uniform_edges(n, lo, hi, seed)
    ug = M.uniform01(seed, np.uint64(1 << 42) + np.arange(n, dtype=np.uint64))
    uk = M.uniform01(seed, np.uint64(1 << 43) + np.arange(n, dtype=np.uint64))
    t = np.cumsum(-np.log1p(-ug) / rate)
    out = []
    for i in range(n):
        q1p0 = uk[i] < q1p0_fraction
        out.append(Request(i, int(L[i]), api.Q1P0 if q1p0 else api.TOTAL_LAGRANGIAN, 0.499 if q1p0 else 0.3,
                           seed * 100003 + i, float(t[i]), isochoric=bool(q1p0) and isochoric))
    return out
