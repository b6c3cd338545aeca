"""Launch-shape autotuner.

Times candidate values of the tuning knobs on the problem actually loaded in
a Solid (device event time of the affected kernels) and leaves the fastest
in place.  The knobs change only launch shapes and launch order, never the
arithmetic of an entry's contributions beyond their summation order, so
every candidate gives a valid result.

Knobs:
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


def _time_stiffness(solid: api.Solid, reps: int) -> float:
    solid.form_lhs()            # warm (builds the launch list)
    solid.synchronize()
    solid.timing_reset()
    for _ in range(reps):
        solid.form_lhs()
    solid.synchronize()
    return solid.timing()["stiffness_ms"] / reps


def tune_stiffness(solid: api.Solid, shapes=None, chunks=None, reps: int = 2) -> dict:
    """Best shape at the default chunk, then the best chunk for that shape.
    Returns {"shape": s, "chunk": c, "ms": {"shape": {...}, "chunk": {...}}}."""
    if shapes is None:
        shapes = list(range(len(api.STIFF_SHAPE_EPB)))
    if chunks is None:
        chunks = stiffness_chunk_candidates(solid.nel)
    t_shape, t_chunk = {}, {}
    solid.timing_enable(True)
    try:
        solid.set_tuning(api.TUNE_STIFF_CHUNK, -1)
        for v in shapes:
            solid.set_tuning(api.TUNE_STIFF_SHAPE, v)
            t_shape[v] = _time_stiffness(solid, reps)
        best_shape = min(t_shape, key=t_shape.get)
        solid.set_tuning(api.TUNE_STIFF_SHAPE, best_shape)
        for v in chunks:
            solid.set_tuning(api.TUNE_STIFF_CHUNK, v)
            t_chunk[v] = _time_stiffness(solid, reps)
        best_chunk = min(t_chunk, key=t_chunk.get) if t_chunk else -1
        solid.set_tuning(api.TUNE_STIFF_CHUNK, best_chunk)
    finally:
        solid.timing_enable(False)
    return {"shape": best_shape, "chunk": best_chunk, "ms": {"shape": t_shape, "chunk": t_chunk}}


def tune_spmv_xcd(solid: api.Solid, run_pcg) -> dict:
    """Time the four PCG SpMV variants (hardware or XCD-contiguous row order,
    plain or branch-free unrolled inner loop); run_pcg() must run a short PCG
    (e.g. one Newton correction with a small iteration cap) from a fixed
    state.  Sets the fastest."""
    times = {}
    solid.timing_enable(True)
    try:
        for v in (0, 1, 2, 3):
            solid.set_tuning(api.TUNE_SPMV_XCD, v)
            run_pcg()
            solid.synchronize()
            solid.timing_reset()
            run_pcg()
            solid.synchronize()
            t = solid.timing()
            times[v] = t["spmv_ms"] / max(1, t["spmv_launches"])
    finally:
        solid.timing_enable(False)
    best = min(times, key=times.get)
    solid.set_tuning(api.TUNE_SPMV_XCD, best)
    return {"xcd": best, "ms": times}
