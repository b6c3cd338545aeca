"""Batches of independent problems (e.g. FE^2 micro-scale RVEs) in one
device context.

The meshes are packed as a disjoint union whose sub-meshes start at node
multiples of 64 (``tg_set_segments``); padding nodes belong to no element and
carry fixed kinematic cards, so they add no equations.  Every sub-problem
keeps its own Newton/PCG control and the same arithmetic as when solved
alone.
"""
from __future__ import annotations

from dataclasses import dataclass

from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import api
from .mesh import HexMesh, KBC


@dataclass
class Problem:
    mesh: HexMesh
    kbc: KBC
    E: list      # Young's modulus per local material id
    nu: list     # Poisson ratio per local material id
    u0: np.ndarray = None  # optional predictor [nnd][3] at the full load (e.g. the affine field u = H X)


@dataclass
class Packed:
    X: np.ndarray
    conn: np.ndarray
    mat: np.ndarray
    kbc: KBC
    node_begin: np.ndarray   # [nseg+1] first node of each sub-problem (padded)
    node_count: np.ndarray   # [nseg] real nodes of each sub-problem
    materials: list          # [(E, nu)] global table
    u0: np.ndarray = None    # [nnd][3] predictor at the full load, or None


def pack(problems: list) -> Packed:
    """The sub-problems one after another in one mesh: each starts on a
    multiple of 64 nodes (a slice of the solver operator), the padding nodes
    fully prescribed to zero; material ids map into one global (E, nu)
    table.  Each array is one concatenation of the problems' own arrays and
    views of shared padding arrays, plus one offset pass."""
    mats, index, gids, gbeg = [], {}, [], [0]
    for p in problems:
        for E, nu in zip(p.E, p.nu):
            key = (float(E), float(nu))
            if key not in index:
                index[key] = len(mats)
                mats.append(key)
            gids.append(index[key])
        gbeg.append(len(gids))
    nnds = np.array([p.mesh.nnd for p in problems], dtype=np.int64)
    pads = (-nnds) % 64
    begins = np.zeros(len(problems) + 1, dtype=np.int64)
    begins[1:] = np.cumsum(nnds + pads)
    nels = np.array([p.mesh.conn.shape[0] for p in problems], dtype=np.int64)
    zeros3 = np.zeros((64, 3))

    def nodal(field):
        return np.concatenate([a for p, q in zip(problems, pads) for a in (field(p), zeros3[:q])])

    def build_conn():
        conn = np.concatenate([p.mesh.conn for p in problems]).astype(np.int32, copy=False)
        conn += np.repeat(begins[:-1].astype(np.int32), nels)[:, None]
        return conn

    def build_mat():
        local = np.concatenate([p.mesh.mat for p in problems]).astype(np.int32, copy=False)
        local += np.repeat(np.asarray(gbeg[:-1], dtype=np.int32), nels)
        return np.asarray(gids, dtype=np.int32)[local]

    def build_kbc():
        # each problem's cards, then its padding nodes' (3 dofs each): node
        # numbers local to the problem's padding start, then one offset pass
        node3 = np.repeat(np.arange(64, dtype=np.int32), 3)
        dof3 = np.tile(np.arange(3, dtype=np.int32), 64)
        zero_v = np.zeros(192)
        nks = np.array([len(p.kbc.nodes) for p in problems], dtype=np.int64)
        kn = np.concatenate([a for p, q in zip(problems, pads)
                             for a in (p.kbc.nodes, node3[:3 * q])]).astype(np.int32, copy=False)
        kd = np.concatenate([a for p, q in zip(problems, pads)
                             for a in (p.kbc.dofs, dof3[:3 * q])]).astype(np.int32, copy=False)
        kv = np.concatenate([a for p, q in zip(problems, pads)
                             for a in (p.kbc.values, zero_v[:3 * q])]).astype(np.float64, copy=False)
        shift = np.empty(2 * len(problems), dtype=np.int32)
        shift[0::2] = begins[:-1]               # a problem's own cards
        shift[1::2] = begins[:-1] + nnds        # its padding nodes
        lens = np.empty(2 * len(problems), dtype=np.int64)
        lens[0::2] = nks
        lens[1::2] = 3 * pads
        kn += np.repeat(shift, lens)
        return KBC(kn, kd, kv)

    jobs = [lambda: nodal(lambda p: p.mesh.X), build_conn, build_mat, build_kbc]
    if any(p.u0 is not None for p in problems):
        jobs.append(lambda: nodal(lambda p: p.u0 if p.u0 is not None else np.zeros_like(p.mesh.X)))
    if int(begins[-1]) >= max(1 << 16, 2048 * len(problems)):
        # large problems: the copies release the interpreter lock, so the
        # arrays are built side by side (many small problems would contend
        # for the lock in the list building instead)
        with ThreadPoolExecutor(len(jobs)) as ex:
            out = [f.result() for f in [ex.submit(j) for j in jobs]]
    else:
        out = [j() for j in jobs]
    X, conn, mat, kbc = out[:4]
    u0 = out[4] if len(out) > 4 else None
    # the last sub-problem needs no trailing padding, but keeping it is harmless
    return Packed(X, conn, mat, kbc, begins.astype(np.int32), nnds.astype(np.int32), mats, u0)


def make_solid(pk: Packed, formulation=api.TOTAL_LAGRANGIAN, precision=api.FP64, device=0,
               precond=api.PC_JACOBI) -> api.Solid:
    """A context for the packed problems (the multigrid preconditioner needs
    a single problem)."""
    return configure(api.Solid(device), pk, formulation, precision, precond)


def configure(s: api.Solid, pk: Packed, formulation=api.TOTAL_LAGRANGIAN, precision=api.FP64,
              precond=api.PC_JACOBI) -> api.Solid:
    """Set up context s, new or used before, for the packed problems.  A used
    context keeps its device allocations where they are large enough, so a
    server that keeps one context per kind of batch allocates nothing once
    its largest batch has been seen; the displacements start from zero as in
    a new context."""
    s.set_mesh(pk.X, pk.conn, pk.mat)
    for m, (E, nu) in enumerate(pk.materials):
        s.set_material(m, *api.mu_kappa(E, nu))
    s.set_formulation(formulation)
    s.set_precision(precision)
    s.set_kbc(pk.kbc.nodes, pk.kbc.dofs, pk.kbc.values)
    s.set_segments(pk.node_begin)
    s.set_preconditioner(precond)
    s.initialize()
    return s


def unpack(pk: Packed, u: np.ndarray) -> list:
    return [u[b:b + n].copy() for b, n in zip(pk.node_begin[:-1], pk.node_count)]


def static_solve_batch(s: api.Solid, pk: Packed, num_steps: int, opts=None):
    """FEManagerT::Solve's static loop for every sub-problem at once; returns
    per-problem converged displacements of the last step and the per-step
    Newton statistics [step][problem].  With predictors (Problem.u0) the
    first step starts from u0 / num_steps instead of zero; the reference
    starts every run from zero, so this changes the Newton path, not the
    converged displacements."""
    stats = []
    if pk.u0 is not None:
        # predictor for the first load step, scaled like the prescribed values
        s.set_displacement(pk.u0 / num_steps)
    for k in range(1, num_steps + 1):
        s.apply_kbc(k / num_steps)
        st = s.newton_batch(opts)
        if any(x.status != 0 for x in st):
            bad = [i for i, x in enumerate(st) if x.status != 0]
            raise api.TahoeGpuError(1, f"Newton failed at step {k} for problems {bad[:10]}")
        s.close_step()
        stats.append(st)
    return unpack(pk, s.displacement()), stats


def solve_one_step_partial(s: api.Solid, pk: Packed, opts=None):
    """One load step to the full load for every sub-problem (from the
    predictor when one is given).  A sub-problem that fails -- an inverted
    element in a residual pass (status 7), divergence or the iteration limit
    -- stops while the others go on to convergence.  Returns per problem its
    displacement (None when it failed) and the Newton statistics."""
    if pk.u0 is not None:
        s.set_displacement(pk.u0)
    s.apply_kbc(1.0)
    st = s.newton_batch(opts)
    u = unpack(pk, s.displacement())
    return [u[i] if st[i].status == 0 else None for i in range(len(st))], st


def solve_adaptive(s: api.Solid, pk: Packed, opts=None, min_dt=1.0 / 1024, log=None):
    """The static load path to t = 1 with step cutting: a step that fails
    (J <= 0 or Newton failure in any sub-problem) is reset and retried at
    half the increment, a converged step doubles the next one
    (FEManagerT's time-step cutting on a failed solve, FEManagerT.cpp:243-266,
    with ResetStep).  Each step starts from the last converged state plus the
    increment of the predictor (Problem.u0) when one is given.  Returns the
    per-problem displacements, the number of converged steps and the Newton
    corrections summed per problem."""
    t, dt, steps = 0.0, 1.0, 0
    newton = np.zeros(len(pk.node_count), dtype=np.int64)
    u = s.displacement()
    while t < 1.0 - 1e-12:
        dt = min(dt, 1.0 - t)
        if pk.u0 is not None:
            s.set_displacement(u + dt * pk.u0)
        s.apply_kbc(t + dt)
        try:
            st = s.newton_batch(opts)
            ok = all(x.status == 0 for x in st)
        except api.TahoeGpuError as e:
            if e.code != 7:
                raise
            ok, st = False, None
        if log:
            log(f"load step t={t:.4g}+{dt:.4g}: {'converged' if ok else 'cut'}")
        if ok:
            s.close_step()
            u = s.displacement()
            t += dt
            steps += 1
            newton += np.array([x.iterations for x in st])
            dt *= 2.0
        else:
            s.reset_step()
            dt *= 0.5
            if dt < min_dt:
                raise api.TahoeGpuError(1, f"load step below {min_dt} at t = {t}")
    return unpack(pk, s.displacement()), steps, newton
