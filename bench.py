#!/usr/bin/env python3
"""Benchmark of the continuum-solid Newton step on one MI355X.

Workload: 128^3 trilinear Q1 hexahedra (2,097,152 elements) in the unit
cube, interior nodes perturbed by U(+-5%) of the element size from seed 2024,
total-Lagrangian compressible neo-Hookean (Simo_isotropic, E = 1000,
nu = 0.3), z = 0 face fixed, z = 1 face u_z = 0.05 t at load step 1 of 5
(t = 0.2), FP64 throughout.  One step = one Newton correction from that
state: residual R, ||R||, stiffness K assembled into the upper MSR matrix,
expansion to the solver's block operator, Jacobi-PCG to ||r|| <= 1e-10 ||R||,
u += du.  Every step restarts from the same state (ResetStep + KBC), so all
steps do identical work.  Mesh, matrix and vectors are resident in HBM
before the timed region starts.

value = elements assembled per second = nel / (device time of the R pass +
device time of the K pass), from HIP events recorded on the solver stream
around every launch inside the timed region.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "elements/s assembled (R+K) + Newton solves/s; p50/p99 RVE latency vs plain FP64"
HBM_PEAK = 8.0e12        # B/s, MI355X HBM3E
FP64_PEAK = 78.6e12      # flop/s, MI355X FP64 vector and matrix (dense)

# algorithmic work per unit (see DESIGN.md "Measurement")
RES_FLOP_PER_EL = 6.2e3          # residual: shape derivatives + kinematics + stress + B^T P
STIFF_FLOP_PER_EL = 83.0e3       # stiffness: 8 x (2*24*6*24 + 2*6*6*24) material + geometric + kinematics


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128, help="elements per cube edge")
    ap.add_argument("--lin-rtol", type=float, default=1e-10)
    ap.add_argument("--cpu-n", type=int, default=64, help="edge of the CPU-baseline sample cube (0: skip)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--profile", action="store_true",
                    help="short run for profilers: 2 Newton corrections with 48 PCG iterations each")
    ap.add_argument("--stiff-chunk", type=int, default=None,
                    help="elements per range of the colour-ordered stiffness launches (default: autotuned)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only the plain copy kernel (1 GiB) used to calibrate HBM counters")
    return ap.parse_args()


def traffic_from_profiles():
    """HBM bytes per SpMV launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    p = os.path.join(ROOT, "profiles", "spmv_bsell_128cube_pmc.md")
    if not os.path.exists(p):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(p).read(), flags=re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(n):
    """The C oracle (single thread) on an n^3 sample of the same workload:
    one residual pass and one stiffness pass."""
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
    return {"value": m.nel / (t2 - t0), "unit": "elements/s", "cores": 1, "kind": "port",
            "sample": f"{n}^3 cube ({m.nel} elements), same generator/seed/BCs, one R + one K pass "
                      f"(R {t1 - t0:.2f} s, K {t2 - t1:.2f} s), C oracle at -O2, 1 thread"}


def main():
    a = parse_args()
    if a.calibrate_copy:
        from tahoe_amd import api
        nbytes = 1 << 30
        ms = api.calib_copy(nbytes, reps=10)
        print(json.dumps({"copy_kernel": "k_copy_calib", "bytes_read": nbytes, "bytes_written": nbytes,
                          "launches": 11, "ms": ms, "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU layout is "
              "described in DESIGN.md and is not measured here.", file=sys.stderr)
        sys.exit(2)
    from tahoe_amd import api, mesh as M

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
    n_eq, msr_len = s.initialize()
    print(f"# setup {time.perf_counter() - t_setup:.1f} s: nel={m.nel} nnd={m.nnd} n_eq={n_eq} "
          f"msr_upper_nnz={msr_len - 1}", file=sys.stderr, flush=True)
    if a.stiff_chunk is not None:
        s.set_tuning(api.TUNE_STIFF_CHUNK, a.stiff_chunk)
        tuned = {"best": a.stiff_chunk, "ms": {}}
    else:
        from tahoe_amd import autotune
        tuned = autotune.tune_stiffness_chunk(s, reps=1)
    print(f"# stiffness chunk {tuned['best']} (autotune ms: {tuned['ms']})", file=sys.stderr, flush=True)
    opts = s.opts(lin_rtol=a.lin_rtol, lin_maxit=48 if a.profile else 200000)
    t_load = 0.2

    def step():
        s.reset_step()
        s.apply_kbc(t_load)
        return s.newton_correction(opts)

    if a.profile:
        s.timing_enable(True)
        for _ in range(2):
            s.reset_step()
            s.apply_kbc(t_load)
            s.newton_correction(opts)
        s.synchronize()
        print(json.dumps({"profile_run": s.timing()}), flush=True)
        return

    for w in range(a.warmup):
        st = step()
        s.synchronize()
        print(f"# warmup {w}: ||R0||={st.error0:.6e} pcg_its={st.lin_iterations}", file=sys.stderr, flush=True)

    s.timing_enable(True)
    s.timing_reset()
    s.synchronize()
    lin_its = []
    t0 = time.perf_counter()
    for k in range(a.steps):
        st = step()
        lin_its.append(st.lin_iterations)
    s.synchronize()
    t1 = time.perf_counter()
    tm = s.timing()
    s.timing_enable(False)
    u = s.displacement()
    assert np.all(np.isfinite(u))

    K = a.steps
    ms_per_step = (t1 - t0) * 1e3 / K
    r_ms = tm["residual_ms"] / K
    k_ms = tm["stiffness_ms"] / K
    value = m.nel / ((r_ms + k_ms) * 1e-3)
    spmv_avg_ms = tm["spmv_ms"] / max(1, tm["spmv_launches"])
    nblocks = (n + 1 + 2 * n) ** 3  # (3n+1)^3 node-neighbour blocks of the structured cube
    vbytes = 8 if a.precision == "fp64" else 4
    spmv_bytes = nblocks * (9 * vbytes + 4) + 2 * 3 * vbytes * m.nnd
    spmv_gbs = spmv_bytes / (spmv_avg_ms * 1e-3) / 1e9
    res_gflops = RES_FLOP_PER_EL * m.nel / (r_ms * 1e-3) / 1e9
    stf_gflops = STIFF_FLOP_PER_EL * m.nel / (k_ms * 1e-3) / 1e9
    its = float(np.mean(lin_its))
    pcg_ms = (tm["spmv_ms"] + tm["vector_ms"]) / K
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
                               "uniaxial 5% load step 1/5, one Newton correction: R + K into MSR + Jacobi-PCG "
                               f"to {a.lin_rtol:g}",
                   "elements": m.nel, "nodes": m.nnd, "equations": n_eq, "msr_upper_nnz": msr_len - 1,
                   "stiffness_chunk": tuned["best"]},
        "roofline": {"bound": "hbm", "achieved": spmv_gbs, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": spmv_gbs / (HBM_PEAK / 1e9), "traffic": traffic_from_profiles(),
                     "kernel": "k_spmv<true> (block sliced-ELL SpMV + p.q)",
                     "algorithmic_bytes_per_launch": spmv_bytes, "avg_launch_ms": spmv_avg_ms},
        "cpu_baseline": cpu_baseline(a.cpu_n) if a.cpu_n > 0 else None,
        "newton_solves_per_s": 1e3 / ms_per_step,
        "residual": {"ms": r_ms, "elements_per_s": m.nel / (r_ms * 1e-3), "gflops": res_gflops,
                     "frac_fp64_peak": res_gflops * 1e9 / FP64_PEAK},
        "stiffness": {"ms": k_ms, "elements_per_s": m.nel / (k_ms * 1e-3), "gflops": stf_gflops,
                      "frac_fp64_peak": stf_gflops * 1e9 / FP64_PEAK},
        "pcg": {"iterations_per_step": its, "ms_per_step": pcg_ms,
                "dof_iterations_per_s": n_eq * its / (pcg_ms * 1e-3) if pcg_ms > 0 else None,
                "expand_ms": tm["expand_ms"] / K},
        "timing_ms": tm,
        "autotune_stiffness_ms": {str(k): v for k, v in tuned["ms"].items()},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
