"""Continuous-batching server for RVE solve requests on one GPU.

Requests (tahoe_amd.rve.Request) arrive over time.  A request is admitted
when the device-memory footprint of everything admitted stays under a
budget; admitted requests of one formulation and size class are grouped
into a mini-batch (at most `max_batch` requests, or whatever has arrived
when a worker is free) that is packed into one device context and solved to
Newton convergence by a worker thread on that context's own HIP stream.
A batch's PCG runs until its slowest member converges, so batches are
kept moderate (max_batch) and of one size class, and several run at once.  Up to
`workers` mini-batches are in flight at once, so requests that arrive while
others run are admitted as memory and workers free up and their work
overlaps on the device.  Each request's latency is completion time minus
arrival time.

Results are deterministic per request: a sub-problem's arithmetic does not
depend on the batch it lands in (segmented reductions, per-problem Newton
control), so the same request gives the same displacement alone or in any
mix.
"""
from __future__ import annotations

import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import api, batch


def footprint_bytes(req, precision=api.MIXED) -> int:
    """Device bytes of one request inside a batch context (upper bound for
    a cube of L^3 hexes): per element the 300-entry MSR slot map, the 72-entry
    operator slot map, connectivity, material and colour lists; per node the
    upper MSR pattern and values (<= 123 entries), 27 operator blocks (column,
    9 FP64 values, 9 FP32 values in mixed precision, 9-entry expansion map)
    and about a dozen node-space vectors."""
    nel = req.L ** 3
    nnd = (req.L + 1) ** 3 + 63          # padding to a 64-node boundary
    per_el = 4 * (300 + 72 + 8 + 1 + 2)
    val_b = 8 + (4 if precision == api.MIXED else 0)
    per_node = 123 * (4 + val_b) + 27 * (4 + 9 * val_b + 9 * 4) + 12 * 3 * 8
    return nel * per_el + nnd * per_node


@dataclass
class Result:
    rid: int
    u: np.ndarray
    newton_iterations: int
    arrival: float
    admitted: float
    finished: float
    batch_id: int
    batch_size: int
    status: int = 0          # 0 converged; otherwise the ExceptionT code / Newton failure of the last attempt
    load_steps: int = 1

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class ServerStats:
    results: dict = field(default_factory=dict)
    batches: list = field(default_factory=list)
    wall: float = 0.0


def precond_for(req) -> int:
    """Inner PCG preconditioner of a served request: the 3x3 nodal block
    Jacobi for the near-incompressible Q1P0 requests (it couples a node's
    dofs, which the volumetric stiffness dominates: about 1.4x fewer PCG
    iterations there), point Jacobi otherwise."""
    return api.PC_BLOCK_JACOBI if req.formulation == api.Q1P0 else api.PC_JACOBI


def size_class(req) -> int:
    """Requests are batched with others of the same formulation and size
    class (edges 4-8, 9-15, 16+), so a batch's members converge in similar
    iteration counts."""
    return 0 if req.L <= 8 else (1 if req.L < 16 else 2)


def solve_request(req, precision, opts, device=0, log=None, precond=None, one_step_failed=False):
    """One request on its own context: one load step from the affine
    predictor, and when that fails the adaptive load path (batch.solve_adaptive).
    Returns (u, Newton corrections, load steps).  A request's result does not
    depend on which batch, if any, it was served in.  precond None: the
    server's choice (precond_for).  one_step_failed: the one-step solve is
    known to fail (it failed inside a batch, whose arithmetic is the same
    alone), so go straight to the adaptive path, which starts from a new
    context either way: the result is the same."""
    pc = precond_for(req) if precond is None else precond
    pk = batch.pack([req.problem()])
    if not one_step_failed:
        s = batch.make_solid(pk, formulation=req.formulation, precision=precision, device=device, precond=pc)
        try:
            try:
                us, st = batch.static_solve_batch(s, pk, 1, opts)
                return us[0], st[-1][0].iterations, 1
            except api.TahoeGpuError:
                pass
        finally:
            s.close()
    s = batch.make_solid(pk, formulation=req.formulation, precision=precision, device=device, precond=pc)
    try:
        us, steps, newton = batch.solve_adaptive(
            s, pk, opts, log=(lambda m: log(f"request {req.rid} (L={req.L}): {m}")) if log else None)
    finally:
        s.close()
    return us[0], int(newton[0]), steps


class RVEServer:
    def __init__(self, precision=api.MIXED, workers=6, max_batch=32, budget_bytes=64 << 30, lin_rtol=None,
                 device=0, keep_u=True, on_batch=None, log=None, reuse_contexts=True):
        self.precision = precision
        # each worker keeps one context per (formulation, size class,
        # preconditioner) alive between batches and sets it up again for the
        # next batch (batch.configure): no allocation once the kind's largest
        # batch has been seen, and no free per batch.  These contexts hold
        # device memory beyond the admission budget, which counts the
        # requests in flight: at most workers x 6 contexts, each the size of
        # the largest batch of its kind so far.
        self.reuse_contexts = reuse_contexts
        self._local = threading.local()
        self.workers = workers
        self.max_batch = max_batch
        self.budget = budget_bytes
        self.device = device
        self.keep_u = keep_u
        self.log = log               # log(message) for load-step cutting of re-solved requests
        self.on_batch = on_batch     # on_batch(batch_id, size, formulation, seconds) after each batch
                                     # (seconds = -1 when the batch failed and is re-solved request by request)
        # inner tolerance: Newton is driven by the FP64 residual, so a looser
        # inner solve costs at most a Newton step, never accuracy (1e-3 is
        # also NLSolver_NK's default linear tolerance)
        rt = lin_rtol if lin_rtol is not None else (1e-3 if precision == api.MIXED else 1e-12)
        self.opts = api.Solid.opts(lin_rtol=rt)
        self._lock = threading.Condition()
        self._queue = deque()        # arrived, not yet admitted
        self._in_use = 0
        self._next_batch = 0
        self._closed = False
        self.stats = ServerStats()

    # -- admission ------------------------------------------------------
    def _take_batch(self):
        """Called with the lock held: pop a mini-batch of one formulation
        whose footprint fits the remaining budget; None if nothing fits."""
        if not self._queue:
            return None
        head = self._queue[0]
        form = head.formulation
        picked, size = [], 0
        for r in list(self._queue):
            if r.formulation != form or size_class(r) != size_class(head):
                continue
            fb = footprint_bytes(r, self.precision)
            if self._in_use + size + fb > self.budget and picked:
                break
            if self._in_use + size + fb > self.budget and self._in_use > 0:
                break  # wait for memory to free up
            picked.append(r)
            size += fb
            if len(picked) >= self.max_batch:
                break
        if not picked:
            return None
        for r in picked:
            self._queue.remove(r)
        self._in_use += size
        bid = self._next_batch
        self._next_batch += 1
        return bid, picked, size

    # -- worker ---------------------------------------------------------
    def _solve_group(self, reqs):
        """Solve reqs as one batch context in one load step (from the affine
        predictor).  Returns per request its displacement (None when that
        sub-problem failed; the others are unaffected) and its Newton
        corrections; raises TahoeGpuError when the batch as a whole fails."""
        pk = batch.pack([r.problem() for r in reqs])
        form, pc = reqs[0].formulation, precond_for(reqs[0])
        if not self.reuse_contexts:
            s = batch.make_solid(pk, formulation=form, precision=self.precision, device=self.device, precond=pc)
            try:
                us, st = batch.solve_one_step_partial(s, pk, self.opts)
            finally:
                s.close()
            return us, [x.iterations for x in st]
        pool = self._pool()
        key = (form, size_class(reqs[0]), pc)
        s = pool.pop(key, None) or api.Solid(self.device)
        try:
            batch.configure(s, pk, formulation=form, precision=self.precision, precond=pc)
            us, st = batch.solve_one_step_partial(s, pk, self.opts)
        except BaseException:
            s.close()   # a context whose batch raised is not used again
            raise
        pool[key] = s
        return us, [x.iterations for x in st]

    def _pool(self) -> dict:
        """This worker thread's live contexts."""
        if not hasattr(self._local, "pool"):
            self._local.pool = {}
        return self._local.pool

    def _close_pool(self):
        for s in self._pool().values():
            s.close()
        self._local.pool = {}

    def _solve(self, bid, reqs, t0):
        """A mini-batch in one load step.  Its converged members finish with
        the batch; a member that failed (an inverted element, divergence)
        goes through solve_request on its own afterwards, and when the batch
        as a whole fails every member does.  A request that cannot be solved
        is reported with its status and u None; a worker never dies on one."""
        t_adm = time.perf_counter() - t0
        done = {}
        batch_ran = False
        try:
            us, its = self._solve_group(reqs)
            batch_ran = True
            t_batch = time.perf_counter() - t0
            for r, u, n in zip(reqs, us, its):
                if u is not None:
                    done[r.rid] = (u, n, 0, 1, t_batch)
        except api.TahoeGpuError:
            pass
        retry = [r for r in reqs if r.rid not in done]
        if retry and self.on_batch:
            self.on_batch(bid, len(retry), reqs[0].formulation, -1.0)  # -1: members re-solved one by one
        for r in retry:
            try:
                # a member that failed in a batch that ran fails the same way
                # alone in one step: straight to the adaptive load path
                u, n, steps = solve_request(r, self.precision, self.opts, self.device, self.log,
                                            one_step_failed=batch_ran)
                done[r.rid] = (u, n, 0, steps, time.perf_counter() - t0)
            except api.TahoeGpuError as e:
                done[r.rid] = (None, 0, e.code, 0, time.perf_counter() - t0)
        out = []
        for r in reqs:
            u, n, code, steps, t_end = done[r.rid]
            out.append(Result(r.rid, u if self.keep_u else None, n, r.arrival, t_adm, t_end, bid, len(reqs), code,
                              steps))
        return out

    def _worker(self, t0):
        try:
            self._serve_batches(t0)
        finally:
            self._close_pool()

    def _serve_batches(self, t0):
        while True:
            with self._lock:
                job = None
                while job is None:
                    job = self._take_batch()
                    if job is None:
                        if self._closed and not self._queue:
                            return
                        self._lock.wait(timeout=0.01)
            bid, reqs, size = job
            try:
                res = self._solve(bid, reqs, t0)
            finally:
                with self._lock:
                    self._in_use -= size
                    self._lock.notify_all()
            with self._lock:
                for x in res:
                    self.stats.results[x.rid] = x
                self.stats.batches.append((bid, len(reqs), reqs[0].formulation, res[0].admitted,
                                           max(x.finished for x in res)))
            if self.on_batch:
                self.on_batch(bid, len(reqs), reqs[0].formulation, max(x.finished for x in res) - res[0].admitted)

    # -- driver ---------------------------------------------------------
    def run(self, requests, realtime=True):
        """Replay `requests` (sorted by arrival, seconds from start) and serve
        them; returns ServerStats with per-request results."""
        api.init_device(self.device)   # runtime start-up here, not in several workers at once
        t0 = time.perf_counter()
        threads = [threading.Thread(target=self._worker, args=(t0,), daemon=True) for _ in range(self.workers)]
        for th in threads:
            th.start()
        for r in sorted(requests, key=lambda x: x.arrival):
            if realtime:
                dt = r.arrival - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
            with self._lock:
                self._queue.append(r)
                self._lock.notify_all()
        with self._lock:
            self._closed = True
            self._lock.notify_all()
        for th in threads:
            th.join()
        self.stats.wall = time.perf_counter() - t0
        return self.stats


def latency_summary(stats: ServerStats) -> dict:
    ok = [r for r in stats.results.values() if r.status == 0]
    lat = np.array([r.latency for r in ok])
    return {"requests": len(stats.results), "converged": len(ok), "failed": len(stats.results) - len(ok),
            "retried_with_load_steps": sum(1 for r in ok if r.load_steps > 1),
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "solves_per_s": len(ok) / stats.wall, "wall_s": stats.wall,
            "batches": len(stats.batches)}
