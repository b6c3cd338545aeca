#!/usr/bin/env python3
"""Benchmark of the continuum-solid Newton step on one MI355X.

Default workload (`cube`): 128^3 trilinear Q1 hexahedra (2,097,152 elements)
in the unit cube, interior nodes perturbed by U(+-5%) of the element size
from seed 2024, total-Lagrangian compressible neo-Hookean (Simo_isotropic,
E = 1000, nu = 0.3), z = 0 face fixed, z = 1 face u_z = 0.05 t at load step
1 of 5 (t = 0.2), FP64 throughout.

One step = one R + K assembly into the inner solver's operator at the state
of the first Newton correction: residual R, ||R||, stiffness K (into the MSR
and expanded, or straight into the block operator), 1/diag.  Mesh, matrix
and vectors are resident in HBM before the timed region starts.
value = elements assembled per second = nel / (device time of the R pass +
device time of the K pass), from HIP events recorded on the solver stream
around every launch inside the timed region.

Beside the timed loop: one complete Newton correction (the same assembly,
Jacobi-PCG to 1e-10, u += du) timed once by wall clock (newton_solves_per_s),
a short event-timed correction for the PCG SpMV's launch duration (the
roofline line), and the C oracle's R + K passes on a 64^3 sample on the
host cores (cpu_baseline).  A watchdog prints the line before --time-budget
seconds even if the long correction is still running.
"""
from __future__ import annotations

import argparse
import collections
import faulthandler
import json
from concurrent.futures import ThreadPoolExecutor
import os
import re
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "elements/s assembled (R+K) + Newton solves/s; p50/p99 RVE latency vs plain FP64"
HBM_PEAK = 8.0e12        # B/s, MI355X HBM3E
FP64_PEAK = 78.6e12      # flop/s, MI355X FP64 vector and matrix (dense)
FP32_MFMA_PEAK = 157.3e12  # flop/s, MI355X FP32 matrix (dense)

# algorithmic work per element (flop counts of the kernels' arithmetic, not of the launch)
RES_FLOP_PER_EL = 6.2e3          # residual: shape derivatives + kinematics + stress + B^T P
STIFF_FLOP_PER_EL = 83.0e3       # stiffness: 8 x (2*24*6*24 + 2*6*6*24) material + geometric + kinematics


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=128, help="elements per cube edge")
    ap.add_argument("--lin-rtol", type=float, default=1e-10)
    ap.add_argument("--cpu-n", type=int, default=64, help="edge of the CPU-baseline sample cube (0: skip)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--precond", choices=["jacobi", "block", "amg"], default="jacobi",
                    help="inner PCG preconditioner: point Jacobi (the reference's DiagonalMatrixT), 3x3 nodal "
                         "block Jacobi, or one V-cycle of aggregation multigrid")
    ap.add_argument("--mixed-n", type=int, default=24, help="cube edge of the mixed-steps workload")
    ap.add_argument("--mixed-lin-rtol", type=float, default=1e-3,
                    help="mixed-steps: relative tolerance of the mixed leg's inner PCG (Newton runs to 1e-10 on the "
                         "FP64 residual)")
    ap.add_argument("--mixed-checkpoints", type=int, default=10,
                    help="mixed-steps: solve the first K of the 10 load-step checkpoints (default all)")
    ap.add_argument("--fp64-lin-rtol", type=float, default=1e-10,
                    help="mixed-steps: relative tolerance of the FP64 leg's inner PCG (the same as --mixed-lin-rtol "
                         "compares the precisions like for like)")
    ap.add_argument("--profile", action="store_true",
                    help="short run for profilers: every SpMV variant timed, the full correction capped at 480 "
                         "PCG iterations, no CPU baseline")
    ap.add_argument("--time-budget", type=float, default=420.0,
                    help="seconds after which the line is printed with what has been measured (0: no limit)")
    ap.add_argument("--autotune", type=int, choices=[0, 1], default=1,
                    help="1: time the launch-shape candidates on the loaded problem first (a few seconds)")
    ap.add_argument("--spmv-sample", type=int, default=256,
                    help="PCG iterations of the event-timed correction that measures the SpMV launch")
    ap.add_argument("--correction", type=int, choices=[0, 1], default=1,
                    help="1: time one complete Newton correction (PCG to --lin-rtol) by wall clock")
    ap.add_argument("--correction-maxit", type=int, default=200000,
                    help="PCG iteration cap of the complete correction")
    ap.add_argument("--pcg-graph", type=int, choices=[0, 1], default=1,
                    help="1: the PCG chunks are replayed as a captured HIP graph (default); 0: direct launches of "
                         "the same kernels (for tracing tools that do not follow graph launches)")
    ap.add_argument("--step-n", type=int, default=128, help="newton-step: cube edge")
    ap.add_argument("--amg-reuse", type=int, default=None,
                    help="newton-step, mixed-steps: tangents of a Newton solve served by one multigrid numeric "
                         "set-up (TG_TUNE_AMG_REUSE; default: the library's, 6)")
    ap.add_argument("--step-legs", default="plain_fp64,fp64_jacobi_1e-3,fp64_multigrid_1e-3,mixed",
                    help="newton-step: the solver paths to run (comma list; the factors need all four)")
    ap.add_argument("--step-stretch", type=float, default=0.005,
                    help="newton-step: top-face displacement of the one load step, fraction of the height")
    ap.add_argument("--workload", choices=["cube", "mixed-steps", "rve-batch", "rve-serve", "newton-step",
                                           "rve-setup"],
                    default="cube",
                    help="cube: one Newton correction on the n^3 cube (default); mixed-steps: the 20%%/10-step "
                         "load path in mixed precision vs FP64, error per step; rve-batch: 4096 RVEs batched; "
                         "rve-serve: Poisson stream of RVE requests; rve-setup: the set-up of rve-batch's "
                         "batches alone")
    ap.add_argument("--rve-count", type=int, default=4096)
    ap.add_argument("--plain-sample", type=int, default=256,
                    help="RVEs of the batch solved one at a time in plain FP64 (rate baseline + error check)")
    ap.add_argument("--batch-elements", type=int, default=1 << 21, help="elements per packed batch")
    ap.add_argument("--batch-lin-rtol", type=float, default=1e-3,
                    help="rve-batch: relative tolerance of the mixed path's inner PCG (Newton runs to 1e-10 on the "
                         "FP64 residual either way)")
    ap.add_argument("--serve-count", type=int, default=512)
    ap.add_argument("--serve-load", type=float, default=0.7, help="offered load as a fraction of capacity")
    ap.add_argument("--serve-raw-h", action="store_true",
                    help="rve-serve: the Q1P0 requests take the raw macro F = I + H (as the served-stream "
                         "configuration states) instead of its isochoric part; many then need cut load steps")
    ap.add_argument("--stiff-chunk", type=int, default=None,
                    help="elements per range of the colour-ordered stiffness launches (default: autotuned)")
    ap.add_argument("--spmv-xcd", type=int, default=None, help="PCG SpMV variant 0..7 (default: autotuned)")
    ap.add_argument("--pcg-fused", type=int, choices=[0, 1], default=None,
                    help="1: PCG direction formed inside the SpMV, 0: its own pass (default: autotuned)")
    ap.add_argument("--assembly", type=int, default=None,
                    help="Newton tangent: 0 into the MSR and expanded, 1 straight into the operator (default: autotuned)")
    ap.add_argument("--stiff-shape", type=int, default=None,
                    help="stiffness workgroup shape 0..15 (default: autotuned)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only the plain copy kernel (1 GiB) used to calibrate HBM counters")
    return ap.parse_args()


PRECOND = {"jacobi": 0, "block": 1, "amg": 2}   # api.PC_JACOBI, PC_BLOCK_JACOBI, PC_AMG
PC_NAME = {"jacobi": "Jacobi", "block": "block-Jacobi", "amg": "multigrid"}


def spmv_kernel_name(mt, fused, variant, pc=0):
    """Template instance of the PCG SpMV that runs (pc: 0 Jacobi, 2 multigrid)."""
    x, u, y = (("true" if variant & b else "false") for b in (1, 2, 4))
    return f"k_fused_spmv<{mt}, {pc}, {x}, {u}, {y}>" if fused else f"k_seg_spmv<{mt}, {x}, {u}, {y}>"


def spmv_operator_bytes(n, nnd, vbytes, symmetric):
    """Algorithmic HBM bytes of the operator part of one SpMV on the
    structured n^3 cube: (3n+1)^3 node-neighbour blocks, nnd of them
    diagonal.  Every block's 9 values and column once; with the symmetric
    read, a lower block's values are its upper partner's, so only the upper
    blocks (diagonal included) bring values and a lower block brings its
    column and the partner's index."""
    nblocks = (n + 1 + 2 * n) ** 3
    if not symmetric:
        return nblocks * (9 * vbytes + 4)
    lower = (nblocks - nnd) // 2
    return (nblocks - lower) * (9 * vbytes + 4) + lower * 8


def traffic_from_profiles(kernel):
    """HBM bytes per launch of `kernel` (a template instance of the 128^3
    FP64 SpMV) from the committed rocprofv3 PMC summary (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration); None when the
    summary does not cover it."""
    p = os.path.join(ROOT, "profiles", "spmv_bsell_128cube_pmc.md")
    if not os.path.exists(p):
        return None
    m = re.search(r"^traffic_bytes_per_launch\[" + re.escape(kernel) + r"\]:\s*([0-9.eE+]+)", open(p).read(),
                  flags=re.M)
    return float(m.group(1)) if m else None


def cpu_threads():
    """Host threads for the multi-threaded CPU baseline: OMP_NUM_THREADS when
    set (the host share of one GPU), else every core."""
    n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(env))) if env.isdigit() and int(env) > 0 else n


def cpu_baseline(n):
    """The C oracle on an n^3 sample of the same workload: one residual pass
    and one stiffness pass, single-threaded in element order (the reference's
    configuration), then over an element colouring on all host threads.
    Returns (single-thread line, all-threads line)."""
    from oracle import oracle as O
    from tahoe_amd import mesh as M
    m = M.cube(n, seed=2024)
    kbc = M.kbc_uniaxial(m, 0.05)
    eq, neq = O.equation_numbers(m.nnd, kbc.nodes, kbc.dofs)
    mu, ka = O.mu_kappa(1000.0, 0.3)
    u = np.zeros((m.nnd, 3))
    u[kbc.nodes, kbc.dofs] = kbc.values * 0.2
    b = O.msr_pattern(neq, m.conn, eq)
    t0 = time.perf_counter()
    O.residual(O.TL, m.X, m.conn, m.mat, mu, ka, u, eq, neq)
    t1 = time.perf_counter()
    O.stiffness(O.TL, m.X, m.conn, m.mat, mu, ka, u, eq, neq, b)
    t2 = time.perf_counter()
    one = {"value": m.nel / (t2 - t0), "unit": "elements/s", "cores": 1, "kind": "port",
           "sample": f"{n}^3 cube ({m.nel} elements), same generator/seed/BCs, one R + one K pass "
                     f"(R {t1 - t0:.2f} s, K {t2 - t1:.2f} s), C oracle at -O2, 1 thread"}
    nt = cpu_threads()
    colors = O.color_elements(m.nnd, m.conn)
    t3 = time.perf_counter()
    O.element_passes_parallel(O.TL, m.X, m.conn, m.mat, mu, ka, u, eq, neq, b, colors, nt, want_K=False)
    t4 = time.perf_counter()
    O.element_passes_parallel(O.TL, m.X, m.conn, m.mat, mu, ka, u, eq, neq, b, colors, nt, want_R=False)
    t5 = time.perf_counter()
    par = {"value": m.nel / (t5 - t3), "unit": "elements/s", "cores": nt, "kind": "port",
           "sample": f"the same {n}^3 sample and passes (R {t4 - t3:.2f} s, K {t5 - t4:.2f} s), elements of "
                     f"each of {len(colors[0]) - 1} colours spread over {nt} OpenMP threads, C oracle at -O2"}
    return one, par


class Deadline:
    """Prints the bench line before the time budget runs out.  The line is
    filled in as the run goes; if the budget expires first (say inside the
    long Newton correction, which runs in the library with the GIL
    released), a watchdog thread prints what has been measured so far,
    marked incomplete, and ends the process."""

    def __init__(self, seconds):
        self.lock = threading.Lock()
        self.line = None
        self.printed = False
        self.t0 = time.perf_counter()
        self.seconds = seconds
        if seconds > 0:
            threading.Thread(target=self._watch, daemon=True).start()

    def left(self):
        return self.seconds - (time.perf_counter() - self.t0)

    def emit(self, line=None):
        with self.lock:
            if self.printed:
                return
            if line is not None:
                self.line = line
            if self.line is not None:
                print(json.dumps(self.line), flush=True)
                self.printed = True

    def _watch(self):
        while self.left() > 0:
            time.sleep(min(1.0, max(0.01, self.left())))
        with self.lock:
            if self.printed:
                return
            if self.line is None:
                print(f"bench.py: time budget of {self.seconds:.0f} s spent before the first measurement",
                      file=sys.stderr, flush=True)
                os._exit(1)
            self.line["incomplete"] = f"time budget of {self.seconds:.0f} s reached; later parts not measured"
            print(json.dumps(self.line), flush=True)
            self.printed = True
        os._exit(0)


def main():
    faulthandler.enable()   # a host-side crash prints every thread's Python stack
    a = parse_args()
    if a.calibrate_copy:
        from tahoe_amd import api
        nbytes = 1 << 30
        ms = api.calib_copy(nbytes, reps=10)
        print(json.dumps({"copy_kernel": "k_copy_calib", "bytes_read": nbytes, "bytes_written": nbytes,
                          "launches": 11, "ms": ms, "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)
        return
    if a.profile:
        a.correction_maxit = min(a.correction_maxit, 480)
        a.cpu_n = 0
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); a multi-GPU partition of the "
              "mesh is not built or measured here.", file=sys.stderr)
        sys.exit(2)
    if a.workload == "mixed-steps":
        return bench_mixed_steps(a)
    if a.workload == "rve-batch":
        return bench_rve_batch(a)
    if a.workload == "rve-setup":
        return bench_rve_setup(a)
    if a.workload == "rve-serve":
        return bench_rve_serve(a)
    if a.workload == "newton-step":
        return bench_newton_step(a)
    return bench_cube(a)


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def bench_cube(a):
    """The n^3 cube (default 128^3).  Timed step: one R + K assembly into the
    inner solver's operator at the state of the first Newton correction of
    load step 1 (R, ||R||, K through the MSR or straight into the block
    operator, 1/diag).  Outside the timed loop: one complete Newton
    correction (the same assembly + PCG to lin_rtol + u += du), timed once
    by wall clock, and a short correction with event timing for the SpMV
    launch duration."""
    from tahoe_amd import api, mesh as M

    dl = Deadline(a.time_budget)
    n = a.n
    t_setup = time.perf_counter()
    m = M.cube(n, seed=2024)
    kbc = M.kbc_uniaxial(m, 0.05)
    s = api.Solid(0)
    s.set_mesh(m.X, m.conn, m.mat)
    s.set_material(0, *api.mu_kappa(1000.0, 0.3))
    s.set_formulation(api.TOTAL_LAGRANGIAN)
    s.set_precision(api.MIXED if a.precision == "mixed" else api.FP64)
    s.set_kbc(kbc.nodes, kbc.dofs, kbc.values)
    s.set_preconditioner(PRECOND[a.precond])
    n_eq, msr_len = s.initialize()
    log(f"setup {time.perf_counter() - t_setup:.1f} s: nel={m.nel} nnd={m.nnd} n_eq={n_eq} "
        f"msr_upper_nnz={msr_len - 1}")
    t_load = 0.2
    short = s.opts(lin_rtol=a.lin_rtol, lin_maxit=32)  # two full PCG chunks: no skipped launches

    def at_state():
        # the state of the first correction of load step 1: last converged
        # u (zero) with the prescribed displacements at t = 0.2
        s.reset_step()
        s.apply_kbc(t_load)

    def short_pcg():
        at_state()
        s.newton_correction(short)

    from tahoe_amd import autotune
    t_tune = time.perf_counter()
    if a.stiff_chunk is not None or a.stiff_shape is not None or a.assembly is not None or not a.autotune:
        s.set_tuning(api.TUNE_STIFF_CHUNK, -1 if a.stiff_chunk is None else a.stiff_chunk)
        if a.stiff_shape is not None:
            s.set_tuning(api.TUNE_STIFF_SHAPE, a.stiff_shape)
        s.set_tuning(api.TUNE_NEWTON_ASSEMBLY, 0 if a.assembly is None else a.assembly)
        tuned = {"assembly": s.tuning(api.TUNE_NEWTON_ASSEMBLY), "shape": s.tuning(api.TUNE_STIFF_SHAPE),
                 "chunk": s.tuning(api.TUNE_STIFF_CHUNK), "ms": {}}
    else:
        tuned = autotune.tune_stiffness(s, reps=1, run_correction=short_pcg)
    log(f"Newton assembly {tuned['assembly']} stiffness shape {tuned['shape']} chunk {tuned['chunk']} "
        f"(autotune ms: {tuned['ms']})")
    explicit = a.spmv_xcd is not None or a.pcg_fused is not None
    tuned_spmv = (autotune.tune_spmv_xcd(s, short_pcg) if a.profile or (a.autotune and not explicit)
                  else {"ms_per_iteration": {}})
    if a.spmv_xcd is not None:
        s.set_tuning(api.TUNE_SPMV_XCD, a.spmv_xcd)
    if a.pcg_fused is not None:
        s.set_tuning(api.TUNE_PCG_FUSED, a.pcg_fused)
    tuned_spmv.update(xcd=s.tuning(api.TUNE_SPMV_XCD), fused=s.tuning(api.TUNE_PCG_FUSED))
    s.set_tuning(api.TUNE_PCG_GRAPH, a.pcg_graph)
    log(f"PCG: SpMV variant {tuned_spmv['xcd']}, fused direction {tuned_spmv['fused']} "
        f"(autotune ms per iteration: {tuned_spmv['ms_per_iteration']}); tuning {time.perf_counter() - t_tune:.1f} s")

    def step():
        s.form_rhs(want=False)      # R and ||R|| (the host reads the norm, as Newton's test does)
        s.form_tangent()            # K into the solver operator + 1/diag

    at_state()
    for w in range(a.warmup):
        step()
    s.synchronize()
    log(f"warmup: {a.warmup} assembly steps")

    s.timing_enable(True)
    s.timing_reset()
    s.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        step()
    s.synchronize()
    t1 = time.perf_counter()
    tm = s.timing()
    s.timing_enable(False)
    K = a.steps
    ms_per_step = (t1 - t0) * 1e3 / K
    r_ms = tm["residual_ms"] / K
    k_ms = tm["stiffness_ms"] / K
    e_ms = tm["expand_ms"] / K
    value = m.nel / ((r_ms + k_ms) * 1e-3)
    log(f"{K} assembly steps: {ms_per_step:.2f} ms per step (R {r_ms:.3f} ms, K {k_ms:.3f} ms, "
        f"operator/diagonal {e_ms:.3f} ms) -> {value / 1e6:.1f} M elements/s")

    # SpMV launch duration: a short correction with every launch event-timed
    at_state()
    s.timing_enable(True)
    s.timing_reset()
    st_s = s.newton_correction(s.opts(lin_rtol=a.lin_rtol, lin_maxit=a.spmv_sample))
    s.synchronize()
    ts = s.timing()
    s.timing_enable(False)
    spmv_avg_ms = ts["spmv_ms"] / max(1, ts["spmv_launches"])
    its_s = max(1, ts["spmv_launches"])
    pcg_it_ms = (ts["spmv_ms"] + ts["vector_ms"] + ts["precond_ms"]) / its_s
    nblocks = (n + 1 + 2 * n) ** 3  # (3n+1)^3 node-neighbour blocks of the structured cube
    vbytes = 8 if a.precision == "fp64" else 4
    variant = s.tuning(api.TUNE_SPMV_XCD)
    symmetric = bool(variant & 4)
    fused = s.tuning(api.TUNE_PCG_FUSED) == 1 and a.precond != "block"
    # operator values + columns once; node vectors once: x and y (FP64); the
    # fused iteration reads p_old and z's sources (r and 1/diag for Jacobi, z
    # for the multigrid) instead of x, and writes the new p besides y
    node_vectors = (5 if a.precond == "jacobi" else 4) if fused else 2
    kname = spmv_kernel_name("double" if a.precision == "fp64" else "float", fused, variant,
                             2 if a.precond == "amg" else 0)
    profiled = n == 128 and a.precision == "fp64" and a.precond == "jacobi"
    # algorithmic bytes of the product on the full (both triangles) block
    # operator: every block's values and column, the node vectors once
    spmv_bytes = spmv_operator_bytes(n, m.nnd, vbytes, False) + node_vectors * 3 * 8 * m.nnd
    spmv_gbs = spmv_bytes / (spmv_avg_ms * 1e-3) / 1e9
    # the symmetric read needs from HBM only the upper half and the lower
    # blocks' indices; the rest is the partner rows' values again, from cache
    unique_bytes = spmv_operator_bytes(n, m.nnd, vbytes, symmetric) + node_vectors * 3 * 8 * m.nnd
    log(f"SpMV {kname}: {spmv_avg_ms:.4f} ms per launch over {ts['spmv_launches']} launches "
        f"-> {spmv_gbs:.0f} GB/s algorithmic")
    res_gflops = RES_FLOP_PER_EL * m.nel / (r_ms * 1e-3) / 1e9
    stf_gflops = STIFF_FLOP_PER_EL * m.nel / (k_ms * 1e-3) / 1e9
    # least HBM bytes of a stiffness pass: every MSR value written once, the
    # element's connectivity and its nodes' X and u read once
    stf_bytes = msr_len * (8 if a.precision == "fp64" else 4) + 32 * m.nel + 48 * m.nnd
    stf_peak = FP64_PEAK if a.precision == "fp64" else FP32_MFMA_PEAK
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32/f64",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 Q1 hex TL neo-Hookean (Simo_isotropic E=1000 nu=0.3), +-5% perturbed, "
                               "uniaxial 5% load step 1/5; step = one R + K assembly into the solver operator at "
                               "the first Newton correction's state; one full Newton correction ("
                               f"{PC_NAME[a.precond]}-PCG to {a.lin_rtol:g}) timed once beside it",
                   "elements": m.nel, "nodes": m.nnd, "equations": n_eq, "msr_upper_nnz": msr_len - 1,
                   "newton_assembly": "operator" if tuned["assembly"] == 1 else "msr+expansion",
                   "stiffness_shape_elements_per_workgroup": api.STIFF_SHAPE_EPB[tuned["shape"]],
                   "stiffness_chunk": tuned["chunk"], "spmv_variant": tuned_spmv["xcd"], "pcg_fused": bool(fused),
                   "precision": a.precision, "preconditioner": a.precond},
        "value_is": "elements / (residual pass + stiffness pass device time), HIP events on the solver stream",
        "roofline": {"bound": "hbm", "achieved": spmv_gbs, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": spmv_gbs / (HBM_PEAK / 1e9),
                     "traffic": traffic_from_profiles(kname) if profiled else None,
                     "kernel": kname + (" (3x3-block sliced-ELL SpMV with the direction update p = z + beta p_old "
                                        "formed inside, + per-slice p.q partials)" if fused else
                                        " (3x3-block sliced-ELL SpMV + per-slice p.q partials)"),
                     "algorithmic_bytes_per_launch": spmv_bytes, "avg_launch_ms": spmv_avg_ms,
                     "symmetric_read": symmetric,
                     "unique_hbm_bytes_per_launch": unique_bytes,
                     "unique_hbm_GBps": unique_bytes / (spmv_avg_ms * 1e-3) / 1e9,
                     "launches_timed": ts["spmv_launches"],
                     "why_this_kernel": "the PCG SpMV is the dominant kernel of a Newton solve (85% of a PCG "
                                        "iteration, ~37k iterations per correction at 128^3)"},
        "roofline_assembly": {
            "stiffness": {"bound": "mfma", "achieved": stf_gflops / 1e3, "peak": stf_peak / 1e12,
                          "unit": "TFLOP/s", "frac": stf_gflops * 1e9 / stf_peak, "avg_pass_ms": k_ms,
                          "flop_per_element": STIFF_FLOP_PER_EL,
                          "hbm_GBps_min_bytes": stf_bytes / (k_ms * 1e-3) / 1e9,
                          "kernel": ("k_stiffness_mfma (16x16x4 MFMA tiles)" if api.STIFF_SHAPE_MFMA[tuned["shape"]]
                                     else "k_stiffness_tl / k_stiffness_db (VALU 3x3 node blocks)"),
                          "note": "dense 24x24 B^T C B count per element; peak: " +
                                  ("FP64 (vector and matrix peaks are equal on MI355X)" if a.precision == "fp64"
                                   else "FP32 matrix cores")},
            "residual": {"bound": "fp64", "achieved": res_gflops / 1e3, "peak": FP64_PEAK / 1e12,
                         "unit": "TFLOP/s", "frac": res_gflops * 1e9 / FP64_PEAK, "avg_pass_ms": r_ms,
                         "flop_per_element": RES_FLOP_PER_EL},
        },
        "elements_per_s_tangent_in_solver_operator": m.nel / ((r_ms + k_ms + e_ms) * 1e-3),
        "assembly_ms": {"residual": r_ms, "stiffness": k_ms, "operator_and_diagonal": e_ms,
                        "wall_per_step": ms_per_step},
        "pcg_iteration_ms": pcg_it_ms,
        "dof_iterations_per_s": n_eq / (pcg_it_ms * 1e-3),
        "timing_ms": tm,
        "autotune_tangent_ms": {k: {str(x): y for x, y in v.items()} for k, v in tuned["ms"].items()},
        "autotune_pcg_ms_per_iteration": tuned_spmv["ms_per_iteration"],
    }
    dl.line = line
    if a.cpu_n > 0:
        one, par = cpu_baseline(a.cpu_n)
        line["cpu_baseline"] = one
        line["cpu_baseline_all_threads"] = par
        log(f"CPU baseline: {one['value']:.0f} elements/s on 1 thread, {par['value']:.0f} on {par['cores']}")
    else:
        line["cpu_baseline"] = None
    if a.correction:
        # one complete Newton correction from the same state, graph-replayed
        # PCG (the default path), by wall clock
        at_state()
        s.synchronize()
        log(f"full Newton correction ({PC_NAME[a.precond]}-PCG to {a.lin_rtol:g}, at most {a.correction_maxit} "
            f"iterations); {dl.left():.0f} s of the time budget left")
        t2 = time.perf_counter()
        st = s.newton_correction(s.opts(lin_rtol=a.lin_rtol, lin_maxit=a.correction_maxit))
        s.synchronize()
        wall = time.perf_counter() - t2
        u = s.displacement()
        assert np.all(np.isfinite(u))
        its = int(st.lin_iterations)
        pcg_s = max(1e-9, wall - (r_ms + k_ms + e_ms) * 1e-3)
        line["newton_solves_per_s"] = 1.0 / wall
        line["newton_correction"] = {
            "wall_s": wall, "pcg_iterations": its, "converged": its < a.correction_maxit,
            "error0": st.error0, "pcg_s": pcg_s, "dof_iterations_per_s": n_eq * its / pcg_s,
            "note": "R, ||R||, K into the operator, PCG (" + ("HIP-graph replayed chunks" if a.pcg_graph else
                                                              "directly launched chunks") + "), u += du; wall clock"}
        log(f"correction: {wall:.2f} s, {its} PCG iterations")
    dl.emit()


def base_line(value, unit, steps, warmup, ms_per_step, dtype, config, higher=True):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": ms_per_step, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": "synthetic", "config": config}


def bench_mixed_steps(a):
    """The 20% stretch in 10 load steps (static Newton to 1e-10 relative
    residual every step, step cutting), solved in FP64 and in mixed precision
    on the same mesh with the same preconditioner; the two legs advance
    checkpoint by checkpoint side by side, and the relative displacement
    error of mixed vs FP64 is reported after every checkpoint.  The inner
    PCG tolerance of each leg is an option (the same for both legs gives the
    like-for-like precision comparison).  A time budget ends the run after
    the last whole checkpoint, and the line says how many were solved."""
    from tahoe_amd import api, mesh as M
    n = a.mixed_n
    dl = Deadline(0)
    t_start = time.perf_counter()
    m = M.cube(n, seed=2024)
    kbc = M.kbc_uniaxial(m, 0.2)
    legs = {}
    rtol = {"fp64": a.fp64_lin_rtol, "mixed": a.mixed_lin_rtol}
    for prec in ("fp64", "mixed"):
        s = api.Solid(0)
        s.set_mesh(m.X, m.conn, m.mat)
        s.set_material(0, *api.mu_kappa(1000.0, 0.3))
        s.set_precision(api.MIXED if prec == "mixed" else api.FP64)
        s.set_kbc(kbc.nodes, kbc.dofs, kbc.values)
        s.set_preconditioner(PRECOND[a.precond])
        n_eq, _ = s.initialize()
        if a.precond == "amg" and a.amg_reuse is not None:
            s.set_tuning(api.TUNE_AMG_REUSE, a.amg_reuse)
        log(f"{prec}: context ready ({n_eq} equations)")
        # large meshes: an inner solve is capped at 15,000 iterations (about a
        # minute with the multigrid at 128^3), so no correction runs unbounded
        opts = s.opts(lin_rtol=rtol[prec], lin_maxit=15000 if n >= 96 else 200000)
        # per-kernel event timing on small meshes only: at 96^3 and above the
        # solve launches millions of kernels, and the line reports wall time
        s.timing_enable(n < 96)
        path = api.LoadPath(s, 10, opts, log=lambda msg, prec=prec: log(f"{prec} {msg}"),
                            log_corrections=n >= 96)
        legs[prec] = {"solid": s, "path": path, "u": None, "newton": [], "substeps": [], "pcg": [], "wall_s": 0.0}
    errs = []
    last = min(a.mixed_checkpoints, 10)
    per_checkpoint_s = []
    for k in range(1, last + 1):
        tk0 = time.perf_counter()
        for prec in ("fp64", "mixed"):
            L = legs[prec]
            t0 = time.perf_counter()
            u, nn, ns, lin = L["path"].advance()
            L["solid"].synchronize()
            L["wall_s"] += time.perf_counter() - t0
            L["u"] = u
            L["newton"].append(nn)
            L["substeps"].append(ns)
            L["pcg"].append(lin)
            log(f"{prec} checkpoint {k}: Newton {nn}, sub-steps {ns}, PCG {lin}, leg total {L['wall_s']:.1f} s")
        ud, um = legs["fp64"]["u"], legs["mixed"]["u"]
        errs.append(float(np.abs(um - ud).max() / np.abs(ud).max()))
        per_checkpoint_s.append(time.perf_counter() - tk0)
        log(f"checkpoint {k}: mixed vs FP64 max relative error {errs[-1]:.3e} "
            f"({time.perf_counter() - t_start:.0f} s since start)")
        left = a.time_budget - (time.perf_counter() - t_start)
        if k < last and left < 1.5 * per_checkpoint_s[-1]:
            log(f"stopping after checkpoint {k}: {left:.0f} s of the time budget left, the last checkpoint "
                f"took {per_checkpoint_s[-1]:.0f} s")
            break
    timing = {}
    for prec in legs:
        legs[prec]["solid"].synchronize()
        timing[prec] = legs[prec]["solid"].timing()
        legs[prec]["solid"].close()
    kk = len(errs)
    tot = lambda t: t["residual_ms"] + t["stiffness_ms"]
    timed = n < 96
    mixed, fp64 = legs["mixed"], legs["fp64"]
    # elements/s assembled (R+K device time) when timed, else elements x Newton
    # corrections per second of wall time for the whole mixed solve
    value = (m.nel * sum(mixed["newton"]) / (tot(timing["mixed"]) * 1e-3) if timed else
             m.nel * sum(mixed["newton"]) / mixed["wall_s"])
    line = base_line(value, "elements/s", kk, 0, mixed["wall_s"] * 1e3 / kk, "f32/f64",
                     {"workload": f"{n}^3 Q1 hex TL neo-Hookean, +-5% perturbed, uniaxial 20% in 10 load steps "
                                  f"({'all' if kk == 10 else f'the first {kk}'} solved; a load step that fails is "
                                  f"cut in half and retried), Newton to 1e-10 on the FP64 residual; preconditioner "
                                  f"{a.precond} in both legs; inner PCG to {a.fp64_lin_rtol:g} (FP64 leg) and "
                                  f"{a.mixed_lin_rtol:g} (mixed leg: FP32 K and operator)",
                      "elements": m.nel, "equations": n_eq, "preconditioner": a.precond,
                      "inner_rtol": {"fp64": a.fp64_lin_rtol, "mixed": a.mixed_lin_rtol}})
    line.update({"value_is": ("elements x Newton corrections / (R + K device time), mixed" if timed else
                              "elements x Newton corrections / wall time of the whole mixed solve"),
                 "checkpoints_solved": kk,
                 "rel_err_vs_fp64_per_step": errs, "max_rel_err": max(errs),
                 "newton_per_step": {"mixed": mixed["newton"], "fp64": fp64["newton"]},
                 "pcg_per_step": {"mixed": mixed["pcg"], "fp64": fp64["pcg"]},
                 "substeps_per_step": {"mixed": mixed["substeps"], "fp64": fp64["substeps"]},
                 "wall_s": {"mixed": mixed["wall_s"], "fp64": fp64["wall_s"]},
                 "speedup_wall": fp64["wall_s"] / mixed["wall_s"],
                 "timing_ms": timing})
    dl.emit(line)


def bench_newton_step(a):
    """One full Newton load step on the n^3 cube (Newton to 1e-10 relative
    residual on the FP64 residual from u = 0, top face raised by
    --step-stretch of the height) on four paths that separate what the
    solver and the precision contribute:
      plain_fp64      FP64 operator, Jacobi-PCG to 1e-10 (the bench line's solver);
      fp64_jacobi_1e-3   FP64, Jacobi-PCG to 1e-3 (inexact Newton);
      fp64_multigrid_1e-3  FP64, multigrid-PCG to 1e-3 (+ preconditioner);
      mixed           FP32 stiffness and operator, multigrid-PCG to 1e-3 (+ precision).
    Wall time of the whole step, elements/s = elements / that time, relative
    displacement error of each path vs plain FP64."""
    from tahoe_amd import api, mesh as M
    n = a.step_n
    m = M.cube(n, seed=2024)
    kbc = M.kbc_uniaxial(m, a.step_stretch)
    res = {}
    # inner solves capped (100,000 Jacobi iterations is about 100 s at 128^3)
    legs = (("plain_fp64", api.FP64, api.PC_JACOBI, 1e-10, 100000),
            ("fp64_jacobi_1e-3", api.FP64, api.PC_JACOBI, 1e-3, 100000),
            ("fp64_multigrid_1e-3", api.FP64, api.PC_AMG, 1e-3, 20000),
            ("mixed", api.MIXED, api.PC_AMG, 1e-3, 20000))
    wanted = set(a.step_legs.split(","))
    for leg, prec, pc, rt, cap in legs:
        if leg not in wanted:
            continue
        s = api.Solid(0)
        s.set_mesh(m.X, m.conn, m.mat)
        s.set_material(0, *api.mu_kappa(1000.0, 0.3))
        s.set_precision(prec)
        s.set_kbc(kbc.nodes, kbc.dofs, kbc.values)
        s.set_preconditioner(pc)
        n_eq, _ = s.initialize()
        s.set_tuning(api.TUNE_PCG_GRAPH, a.pcg_graph)
        if pc == api.PC_AMG and a.amg_reuse is not None:
            s.set_tuning(api.TUNE_AMG_REUSE, a.amg_reuse)
        log(f"{leg}: context ready ({n_eq} equations)")
        s.apply_kbc(1.0)
        s.synchronize()
        t0 = time.perf_counter()
        st = s.newton_stepwise(
            s.opts(lin_rtol=rt, lin_maxit=cap),
            log=lambda i, rel, pcg, leg=leg, t0=t0: log(
                f"{leg} correction {i}: |R|/|R0| {rel:.3e}, PCG {pcg}, {time.perf_counter() - t0:.1f} s"))
        s.synchronize()
        wall = time.perf_counter() - t0
        if st.status != 0:
            raise RuntimeError(f"{leg}: Newton did not converge ({st.iterations} corrections)")
        res[leg] = {"u": s.displacement(), "wall_s": wall, "newton": st.iterations, "pcg": st.lin_iterations}
        s.close()
    if len(res) < len(legs):
        # a subset of the paths (e.g. one under a profiler): timings only
        print(json.dumps({leg: {k: v for k, v in r.items() if k != "u"} for leg, r in res.items()}), flush=True)
        return
    ud = res["plain_fp64"]["u"]
    for leg in res:
        res[leg]["rel_err_vs_plain_fp64"] = float(np.abs(res[leg]["u"] - ud).max() / np.abs(ud).max())
        res[leg]["elements_per_s"] = m.nel / res[leg]["wall_s"]
    w = {leg: res[leg]["wall_s"] for leg in res}
    value = m.nel / w["mixed"]
    line = base_line(value, "elements/s", 1, 0, w["mixed"] * 1e3, "f32/f64",
                     {"workload": f"{n}^3 Q1 hex TL neo-Hookean (Simo_isotropic E=1000 nu=0.3), +-5% perturbed, one "
                                  f"load step: top face raised by {a.step_stretch:g} of the height from u = 0, Newton "
                                  "to 1e-10 relative residual on the FP64 residual; four solver paths (see legs)",
                      "elements": m.nel})
    line.update({"value_is": "elements / wall time of the whole Newton step, mixed path",
                 "legs": {leg: {k: v for k, v in res[leg].items() if k != "u"} for leg in res},
                 "speedup_vs_plain_fp64": w["plain_fp64"] / w["mixed"],
                 "contributions": {
                     "inexact_newton (Jacobi 1e-10 -> 1e-3)": w["plain_fp64"] / w["fp64_jacobi_1e-3"],
                     "multigrid (Jacobi -> multigrid, both 1e-3, FP64)": w["fp64_jacobi_1e-3"] / w["fp64_multigrid_1e-3"],
                     "precision (FP64 -> mixed, both multigrid 1e-3)": w["fp64_multigrid_1e-3"] / w["mixed"]},
                 "rel_err_vs_fp64": res["mixed"]["rel_err_vs_plain_fp64"]})
    print(json.dumps(line), flush=True)


def solve_alone(req, precision, opts, problem=None, ctx=None):
    """One request set up and solved alone (the plain one-problem-at-a-time
    path), in context ctx (set up again for it) or a new one.  Returns (u,
    Newton stats, seconds for pack + set-up + solve, seconds for the solve
    alone)."""
    from tahoe_amd import api, batch
    p = problem if problem is not None else req.problem()
    t0 = time.perf_counter()
    pk = batch.pack([p])
    s = batch.configure(ctx if ctx is not None else api.Solid(0), pk, formulation=req.formulation,
                        precision=precision)
    t1 = time.perf_counter()
    us, st = batch.static_solve_batch(s, pk, 1, opts)
    s.synchronize()
    t2 = time.perf_counter()
    if ctx is None:
        s.close()
    return us[0], st[0][0], t2 - t0, t2 - t1


def size_batches(reqs, max_el):
    """Requests sorted by size, cut into batches of at most max_el elements
    (similar sizes together, so a batch's PCG runs about as long as its
    members need)."""
    out, cur, nel = [], [], 0
    for r in sorted(reqs, key=lambda r: (r.L, r.rid)):
        if cur and nel + r.elements > max_el:
            out.append(cur)
            cur, nel = [], 0
        cur.append(r)
        nel += r.elements
    if cur:
        out.append(cur)
    return out


def bench_rve_batch(a):
    """4096 two-phase RVE cubes (edge log-uniform 4-24), affine boundary
    displacement, each solved to Newton convergence (rel 1e-10): batched
    mixed precision vs the plain FP64 path (one problem per context, one at a
    time) on a sample of the same requests."""
    from tahoe_amd import api, batch, rve
    reqs = rve.batch_requests(a.rve_count)
    tot_el = sum(r.elements for r in reqs)
    fp64 = api.Solid.opts(lin_rtol=1e-10)
    mixed = api.Solid.opts(lin_rtol=a.batch_lin_rtol)
    # plain FP64, one at a time, on every k-th request
    k = max(1, len(reqs) // max(1, a.plain_sample))
    sample = reqs[::k][:a.plain_sample]
    # one context, set up again for every RVE (as the batched legs reuse one
    # for every batch): the comparison is of batching and precision, not of
    # context creation
    plain_ctx = api.Solid(0)
    solve_alone(sample[0], api.FP64, fp64, ctx=plain_ctx)  # warm-up (library load)
    plain_u, plain_wall, plain_solve, plain_el = {}, 0.0, 0.0, 0
    probs = {r.rid: r.problem() for r in reqs}   # mesh generation: the client's, not timed
    for i, r in enumerate(sample):
        if i % 32 == 0:
            print(f"# plain FP64 RVE {i}/{len(sample)}", file=sys.stderr, flush=True)
        u, st, wall, solve = solve_alone(r, api.FP64, fp64, probs[r.rid], ctx=plain_ctx)
        assert st.status == 0
        plain_u[r.rid] = u
        plain_wall += wall
        plain_solve += solve
        plain_el += r.elements
    plain_ctx.close()
    print(f"# plain FP64: {len(sample)} RVEs, {plain_el} elements in {plain_wall:.2f} s", file=sys.stderr,
          flush=True)
    # batched over every request, three ways: FP64 with the plain inner
    # tolerance (what batching alone gives), FP64 with the inexact inner
    # tolerance (the solver), mixed precision with it (the precision)
    batches = size_batches(reqs, a.batch_elements)

    def run_batched(tag, precision, opts):
        t0 = time.perf_counter()
        solve_s, errs, newton_max = 0.0, [], 0
        s = api.Solid(0)   # one context, set up again for every batch
        # the next batch is packed on a host thread while this one solves (the
        # library calls release the interpreter lock)
        pool = ThreadPoolExecutor(1)
        nxt = pool.submit(batch.pack, [probs[r.rid] for r in batches[0]])
        for i, b in enumerate(batches):
            pk = nxt.result()
            if i + 1 < len(batches):
                nxt = pool.submit(batch.pack, [probs[r.rid] for r in batches[i + 1]])
            batch.configure(s, pk, formulation=api.TOTAL_LAGRANGIAN, precision=precision)
            t1 = time.perf_counter()
            us, st = batch.static_solve_batch(s, pk, 1, opts)
            s.synchronize()
            solve_s += time.perf_counter() - t1
            newton_max = max(newton_max, max(x.iterations for x in st[0]))
            for r, u in zip(b, us):
                if r.rid in plain_u:
                    ref = plain_u[r.rid]
                    errs.append(float(np.abs(u - ref).max() / max(np.abs(ref).max(), 1e-300)))
        pool.shutdown()
        s.close()
        wall = time.perf_counter() - t0
        log(f"{tag}: {len(batches)} batches in {wall:.2f} s (Newton alone {solve_s:.2f} s)")
        return {"wall_s": wall, "solve_s": solve_s, "elements_per_s": tot_el / wall,
                "solve_only_elements_per_s": tot_el / solve_s,
                "max_rel_err_vs_fp64": max(errs) if errs else None, "compared_rves": len(errs),
                "newton_max": newton_max}

    legs = {"batched_fp64_1e-10": run_batched("batched FP64, PCG to 1e-10", api.FP64, fp64),
            f"batched_fp64_{a.batch_lin_rtol:g}": run_batched(f"batched FP64, PCG to {a.batch_lin_rtol:g}", api.FP64,
                                                              mixed),
            f"batched_mixed_{a.batch_lin_rtol:g}": run_batched(f"batched mixed, PCG to {a.batch_lin_rtol:g}",
                                                               api.MIXED, mixed)}
    main_leg = legs[f"batched_mixed_{a.batch_lin_rtol:g}"]
    wall = main_leg["wall_s"]
    value = tot_el / wall
    plain_rate = plain_el / plain_wall
    r64 = legs["batched_fp64_1e-10"]["elements_per_s"]
    r64i = legs[f"batched_fp64_{a.batch_lin_rtol:g}"]["elements_per_s"]
    line = base_line(value, "elements/s", len(batches), 1, wall * 1e3 / len(batches), "f32/f64",
                     {"workload": f"{len(reqs)} two-phase RVE cubes, edge log-uniform 4-24 (seed 4242), "
                                  "Bernoulli(0.3) phase 2 at 10x stiffness, affine u = H X, |H_ij| <= 0.02, "
                                  f"Newton to 1e-10 per RVE; batched mixed precision (FP32 K + FP32 PCG to "
                                  f"{a.batch_lin_rtol:g})",
                      "rves": len(reqs), "elements": tot_el, "batches": len(batches),
                      "batch_elements": a.batch_elements})
    line.update({
        "rves_per_s": len(reqs) / wall,
        "solve_only_elements_per_s": main_leg["solve_only_elements_per_s"],
        "plain_fp64": {"rves": len(sample), "elements": plain_el, "wall_s": plain_wall,
                       "elements_per_s": plain_rate, "solve_only_elements_per_s": plain_el / plain_solve,
                       "note": "every k-th request, one at a time (one context set up again per RVE), FP64 K + "
                               "Jacobi-PCG to 1e-10"},
        "legs": legs,
        "contributions": {"batching (one at a time -> batched, FP64 PCG to 1e-10)": r64 / plain_rate,
                          f"inexact inner solve (1e-10 -> {a.batch_lin_rtol:g}, batched FP64)": r64i / r64,
                          "precision (FP64 -> mixed, batched, same tolerance)": value / r64i},
        "timed": "every path: pack + set-up of one reused context (tg_set_mesh ... tg_initialize) + Newton to "
                 "convergence (batched: the next batch packed on a host thread during the current solve); mesh "
                 "generation excluded; solve_only: Newton alone",
        "speedup_solve_only": main_leg["solve_only_elements_per_s"] / (plain_el / plain_solve),
        "speedup_vs_plain_fp64": value / plain_rate,
        "max_rel_err_vs_fp64": main_leg["max_rel_err_vs_fp64"], "compared_rves": main_leg["compared_rves"],
        "newton_max": main_leg["newton_max"],
    })
    print(json.dumps(line), flush=True)


def bench_rve_setup(a):
    """The set-up of rve-batch's batches (the 4096 RVE cubes cut into batches
    of at most --batch-elements elements) on one reused context, mixed
    precision, no solve: pack, then tg_set_mesh ... tg_initialize; twice
    over the batches, the second pass reported (the context's buffers then
    have their size)."""
    from tahoe_amd import api, batch, rve
    reqs = rve.batch_requests(a.rve_count)
    probs = {r.rid: r.problem() for r in reqs}
    batches = size_batches(reqs, a.batch_elements)
    s = api.Solid(0)
    rows = []
    for rep in range(2):
        rows = []
        for b in batches:
            t0 = time.perf_counter()
            pk = batch.pack([probs[r.rid] for r in b])
            t1 = time.perf_counter()
            batch.configure(s, pk, formulation=api.TOTAL_LAGRANGIAN, precision=api.MIXED)
            s.synchronize()
            t2 = time.perf_counter()
            rows.append({"rves": len(b), "elements": int(sum(r.elements for r in b)),
                         "pack_ms": (t1 - t0) * 1e3, "setup_ms": (t2 - t1) * 1e3})
    s.close()
    tot_el = sum(r["elements"] for r in rows)
    setup_s = sum(r["setup_ms"] for r in rows) / 1e3
    line = base_line(tot_el / setup_s, "elements/s", len(batches), len(batches), setup_s * 1e3 / len(batches),
                     "f32/f64", {"workload": f"set-up of {len(reqs)} RVE cubes (edge log-uniform 4-24, seed 4242) "
                                             f"in {len(batches)} batches of at most {a.batch_elements} elements, "
                                             "one reused context, mixed precision",
                                 "rves": len(reqs), "elements": tot_el, "batches": len(batches)})
    line.update({"metric": "elements/s set up (tg_set_mesh ... tg_initialize)", "per_batch": rows,
                 "pack_elements_per_s": tot_el / (sum(r["pack_ms"] for r in rows) / 1e3)})
    print(json.dumps(line), flush=True)


def bench_rve_serve(a):
    """Served stream: Poisson arrivals of RVE requests (edge log-uniform
    4-32, half Q1 nu=0.3, half Q1P0 nu=0.499), continuous batching with HBM
    admission; p50/p99 latency and solves/s, next to the plain FP64 path
    (one request at a time, in arrival order) on the same stream."""
    from tahoe_amd import api, rve, serve
    # capacity: a burst of requests all arriving at once (the server's
    # throughput at saturation)
    iso = not a.serve_raw_h
    cal = rve.stream_requests(min(256, a.serve_count), rate=1e9, seed=555, isochoric=iso)
    def progress(bid, size, form, sec):
        what = "failed, re-solving one by one" if sec < 0 else f"in {sec:.2f} s"
        print(f"# batch {bid}: {size} requests ({'Q1P0' if form == api.Q1P0 else 'Q1'}) {what}",
              file=sys.stderr, flush=True)

    def log(msg):
        print(f"# {msg}", file=sys.stderr, flush=True)

    # a request arrives with its mesh: the problems are built before the clock starts
    for r in cal:
        r.problem()
    srv = serve.RVEServer(precision=api.MIXED, keep_u=False, on_batch=progress, log=log)
    st = srv.run(cal, realtime=False)
    cap = len(cal) / st.wall
    cap_els = sum(r.elements for r in cal) / st.wall
    rate = a.serve_load * cap
    print(f"# capacity {cap:.1f} solves/s -> offered {rate:.1f}/s", file=sys.stderr, flush=True)
    reqs = rve.stream_requests(a.serve_count, rate=rate, isochoric=iso)
    for r in reqs:
        r.problem()
    srv = serve.RVEServer(precision=api.MIXED, on_batch=progress, log=log)
    st = srv.run(reqs, realtime=True)
    lat = serve.latency_summary(st)
    # plain FP64 on the same arrivals: one request at a time
    fp64 = api.Solid.opts(lin_rtol=1e-10)
    n_plain = min(len(reqs), a.plain_sample)
    t0 = time.perf_counter()
    clock, plain_lat, errs = 0.0, [], []
    for i, r in enumerate(reqs[:n_plain]):
        if i % 32 == 0:
            print(f"# plain FP64 request {i}/{n_plain}", file=sys.stderr, flush=True)
        t_r = time.perf_counter()
        u, _, _ = serve.solve_request(r, api.FP64, fp64, log=log, precond=api.PC_JACOBI)
        wall = time.perf_counter() - t_r
        clock = max(clock, r.arrival) + wall
        plain_lat.append(clock - r.arrival)
        ref = st.results[r.rid].u
        if ref is not None:
            errs.append(float(np.abs(ref - u).max() / max(np.abs(u).max(), 1e-300)))
    plain_wall = time.perf_counter() - t0
    pl = np.array(plain_lat)
    els = sum(r.elements for r in reqs)
    plain_els = sum(r.elements for r in reqs[:n_plain])
    line = base_line(cap_els, "elements/s", len(reqs), 0, st.wall * 1e3 / len(reqs), "f32/f64",
                     {"workload": f"{len(reqs)} RVE requests, Poisson arrivals at {rate:.1f}/s "
                                  f"({a.serve_load:g} of measured capacity), edge log-uniform 4-32, 50% Q1 nu=0.3 "
                                  "/ 50% Q1P0 nu=0.499 (" + ("isochoric part of the macro F" if iso else
                                                            "raw macro F = I + H") +
                                  "), two-phase ratio 10, |H_ij| <= 0.02, "
                                  "mixed precision (FP32 K, FP64 R and Newton; PCG to 1e-3, block Jacobi for Q1P0), "
                                  "Newton to 1e-10; meshes built before arrival",
                      "requests": len(reqs), "elements": els})
    line.update({"value_is": "elements/s at saturation: a burst of "
                             f"{len(cal)} requests of the same mix served at once",
                 "stream_elements_per_s": els / st.wall, "latency": lat, "capacity_solves_per_s": cap,
                 "plain_fp64": {"requests": n_plain, "p50_s": float(np.percentile(pl, 50)),
                                "p99_s": float(np.percentile(pl, 99)), "wall_s": plain_wall,
                                "solves_per_s": n_plain / plain_wall, "elements_per_s": plain_els / plain_wall,
                                "note": "the first requests of the same stream at the same arrival times, one "
                                        "request at a time on its own context, FP64 throughout, Jacobi-PCG to "
                                        "1e-10"},
                 "speedup_vs_plain_fp64": cap_els / (plain_els / plain_wall),
                 "stream_speedup_vs_plain_fp64": (els / st.wall) / (plain_els / plain_wall),
                 "max_rel_err_vs_fp64": max(errs) if errs else None,
                 "load_steps": {str(k): v for k, v in sorted(collections.Counter(
                     x.load_steps for x in st.results.values()).items())}})
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
