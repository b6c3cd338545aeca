"""Host-side pieces of bench.py: the committed SpMV traffic profile covers
every kernel variant the bench line can run (FP64, Jacobi), so its roofline
`traffic` is never left empty; kernel names, thread counts and the
multi-threaded CPU baseline are well formed."""
import os
import sys

import pytest

ROOT = os.path.join(os.path.dirname(__file__), "..")
sys.path.insert(0, ROOT)

import bench  # noqa: E402


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("variant", range(8))
def test_profile_covers_every_spmv_variant(fused, variant):
    name = bench.spmv_kernel_name("double", fused, variant)
    t = bench.traffic_from_profiles(name)
    assert t is not None, name
    # at least the algorithmic bytes of one launch (the per-node vectors of
    # the fused form included), and not absurdly more
    nnd = 129 ** 3
    alg = bench.spmv_operator_bytes(128, nnd, 8, False) + (5 if fused else 2) * 24 * nnd
    assert alg == 4440104572 + (3 * 3 * 8 * nnd if fused else 0)
    assert alg * 0.95 <= t <= alg * 2.0, (name, t)


def test_kernel_names_and_threads():
    assert bench.spmv_kernel_name("float", False, 3) == "k_seg_spmv<float, true, true, false>"
    assert bench.spmv_kernel_name("double", True, 1, pc=2) == "k_fused_spmv<double, 2, true, false, false>"
    assert bench.spmv_kernel_name("double", False, 5) == "k_seg_spmv<double, true, false, true>"
    assert bench.traffic_from_profiles("k_seg_spmv<float, false, false, false>") is None
    assert 1 <= bench.cpu_threads() <= (os.cpu_count() or 1)


def test_symmetric_operator_bytes():
    """On the n^3 cube: (3n+1)^3 blocks, (n+1)^3 of them diagonal; the
    symmetric read brings the values of the upper half and the diagonal, and
    8 index bytes per lower block."""
    n = 4
    nnd = (n + 1) ** 3
    nb = (3 * n + 1) ** 3
    full = bench.spmv_operator_bytes(n, nnd, 8, False)
    sym = bench.spmv_operator_bytes(n, nnd, 8, True)
    assert full == nb * 76
    lower = (nb - nnd) // 2
    assert (nb - nnd) % 2 == 0 and sym == (nb - lower) * 76 + lower * 8
    assert 0.5 < sym / full < 0.6


def test_cpu_baselines_small_sample():
    one, par = bench.cpu_baseline(8)
    assert one["cores"] == 1 and one["kind"] == "port" and one["value"] > 0
    assert par["cores"] == bench.cpu_threads() and par["value"] > 0
    assert "8^3" in one["sample"] and "colours" in par["sample"]


def test_deadline_prints_partial_line_once(tmp_path):
    """The watchdog prints the line measured so far, marked incomplete, when
    the time budget runs out inside a long call, and the process ends with
    status 0; a line printed in time is printed once."""
    import json
    import subprocess
    code = ("import time, bench\n"
            "d = bench.Deadline(1.0)\n"
            "d.line = {'value': 1.0}\n"
            "time.sleep(30)\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = [json.loads(x) for x in r.stdout.splitlines() if x.startswith("{")]
    assert len(out) == 1 and out[0]["value"] == 1.0 and "incomplete" in out[0]
    code = ("import time, bench\n"
            "d = bench.Deadline(2.0)\n"
            "d.emit({'value': 2.0})\n"
            "time.sleep(3)\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=60)
    out = [json.loads(x) for x in r.stdout.splitlines() if x.startswith("{")]
    assert r.returncode == 0 and out == [{"value": 2.0}]
