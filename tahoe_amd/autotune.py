"""Launch-shape autotuner.

Times candidate values of the tuning knobs on the problem actually loaded in
a Solid (device event time of the affected kernels) and leaves the fastest
in place.  The knobs change only launch shapes and launch order, never the
arithmetic of an entry's contributions beyond their summation order, so
every candidate gives a valid result.

Knobs:
  TUNE_NEWTON_ASSEMBLY -- Newton's tangent straight into the solver's block
  operator (576 scattered values per element, component-major) or into the
  MSR (300 values, row-contiguous clusters) and then expanded; the traffic of
  the two differs by the cache-line reuse of the scatter.
  TUNE_STIFF_SHAPE -- elements per workgroup of the stiffness kernel
  (api.STIFF_SHAPE_EPB): fewer elements per workgroup means more workgroups
  resident per CU, which hides barrier and scatter latency, at the cost of
  idle lanes in the last wave.
  TUNE_STIFF_CHUNK -- elements per range of the colour-ordered stiffness
  launches (0 = one launch per colour).  A range small enough that the
  matrix rows it touches stay in the memory-side cache across the colour
  passes cuts HBM traffic of the scatter; too small a range under-fills the
  GPU.
"""
from __future__ import annotations

from . import api


def stiffness_chunk_candidates(nel: int) -> list:
    out = [0]
    c = 65536
    while c < nel:
        out.append(c)
        c *= 2
    return out


def _time_stiffness(solid: api.Solid, reps: int, run_correction=None) -> float:
    """Mean device time of assembling the Newton tangent: the stiffness pass
    plus what makes it the solver's operator (expansion, inverse diagonal).
    With run_correction (a short Newton correction from a fixed state) the
    Newton path is timed; otherwise tg_form_lhs (the MSR assembly)."""
    step = run_correction or solid.form_lhs
    step()                      # warm (builds launch lists, captures graphs)
    solid.synchronize()
    solid.timing_reset()
    for _ in range(reps):
        step()
    solid.synchronize()
    t = solid.timing()
    return (t["stiffness_ms"] + (t["expand_ms"] if run_correction else 0.0)) / reps


def tune_stiffness(solid: api.Solid, shapes=None, chunks=None, reps: int = 2, run_correction=None) -> dict:
    """Best (Newton assembly path, shape) at the default chunk, then the best
    chunk.  With run_correction the Newton tangent is timed on both paths --
    straight into the block operator, or into the MSR and expanded -- and
    the faster is kept (TUNE_NEWTON_ASSEMBLY); without it the MSR assembly of
    tg_form_lhs is tuned.  Returns {"assembly", "shape", "chunk", "ms"}."""
    if shapes is None:
        shapes = list(range(len(api.STIFF_SHAPE_EPB)))
    if chunks is None:
        chunks = stiffness_chunk_candidates(solid.nel)
    modes = (0, 1) if run_correction else (solid.tuning(api.TUNE_NEWTON_ASSEMBLY),)
    t_shape, t_chunk = {}, {}
    solid.timing_enable(True)
    try:
        solid.set_tuning(api.TUNE_STIFF_CHUNK, -1)
        for m in modes:
            solid.set_tuning(api.TUNE_NEWTON_ASSEMBLY, m)
            for v in shapes:
                solid.set_tuning(api.TUNE_STIFF_SHAPE, v)
                t_shape[(m, v)] = _time_stiffness(solid, reps, run_correction)
        best_mode, best_shape = min(t_shape, key=t_shape.get)
        solid.set_tuning(api.TUNE_NEWTON_ASSEMBLY, best_mode)
        solid.set_tuning(api.TUNE_STIFF_SHAPE, best_shape)
        for v in chunks:
            solid.set_tuning(api.TUNE_STIFF_CHUNK, v)
            t_chunk[v] = _time_stiffness(solid, reps, run_correction)
        best_chunk = min(t_chunk, key=t_chunk.get) if t_chunk else -1
        solid.set_tuning(api.TUNE_STIFF_CHUNK, best_chunk)
    finally:
        solid.timing_enable(False)
    return {"assembly": best_mode, "shape": best_shape, "chunk": best_chunk,
            "ms": {"assembly_shape": {f"{m}/{v}": t for (m, v), t in t_shape.items()}, "chunk": t_chunk}}


def tune_spmv_xcd(solid: api.Solid, run_pcg, fused=(1, 0), reps: int = 2, variants=range(8)) -> dict:
    """Time the PCG iteration over the SpMV variants (hardware or
    XCD-contiguous row order, plain or branch-free unrolled inner loop, lower
    blocks read directly or as their upper partners' transposes), with
    the direction formed inside the SpMV (fused) and as its own pass; run_pcg()
    must run a short PCG (e.g. one Newton correction with a small iteration
    cap) from a fixed state.  The figure compared is device time per
    iteration (SpMV + vector kernels + preconditioner), the best of `reps`
    timed runs: the fused SpMV reads more per launch and saves the direction
    pass.  Sets the fastest."""
    times = {}
    solid.timing_enable(True)
    try:
        for f in fused:
            solid.set_tuning(api.TUNE_PCG_FUSED, f)
            for v in variants:
                solid.set_tuning(api.TUNE_SPMV_XCD, v)
                run_pcg()
                solid.synchronize()
                best = float("inf")
                for _ in range(reps):
                    solid.timing_reset()
                    run_pcg()
                    solid.synchronize()
                    t = solid.timing()
                    its = max(1, t["spmv_launches"])
                    best = min(best, (t["spmv_ms"] + t["vector_ms"] + t["precond_ms"]) / its)
                times[(f, v)] = best
    finally:
        solid.timing_enable(False)
    bf, bv = min(times, key=times.get)
    solid.set_tuning(api.TUNE_PCG_FUSED, bf)
    solid.set_tuning(api.TUNE_SPMV_XCD, bv)
    return {"xcd": bv, "fused": bf, "ms_per_iteration": {f"{f}/{v}": t for (f, v), t in times.items()}}
