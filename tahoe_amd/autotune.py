"""Launch-shape autotuner.

Times candidate values of a tuning knob on the problem actually loaded in a
Solid (device event time of the affected kernels) and leaves the fastest in
place.  The knob values change only the launch order, never the arithmetic
of an entry's contributions beyond their summation order, so every
candidate gives a valid result.

Knobs:
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


def tune_stiffness_chunk(solid: api.Solid, candidates=None, reps: int = 2) -> dict:
    """Time reps stiffness passes (tg_form_lhs) per candidate; set the best.
    Returns {"best": value, "ms": {value: mean ms}}."""
    if candidates is None:
        candidates = stiffness_chunk_candidates(solid.nel)
    was = solid.tuning(api.TUNE_STIFF_CHUNK)
    times = {}
    solid.timing_enable(True)
    try:
        for v in candidates:
            solid.set_tuning(api.TUNE_STIFF_CHUNK, v)
            solid.form_lhs()            # warm (builds the launch list)
            solid.synchronize()
            solid.timing_reset()
            for _ in range(reps):
                solid.form_lhs()
            solid.synchronize()
            times[v] = solid.timing()["stiffness_ms"] / reps
    finally:
        solid.timing_enable(False)
    if not times:
        solid.set_tuning(api.TUNE_STIFF_CHUNK, was)
        return {"best": was, "ms": {}}
    best = min(times, key=times.get)
    solid.set_tuning(api.TUNE_STIFF_CHUNK, best)
    return {"best": best, "ms": times}
