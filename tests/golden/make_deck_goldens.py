#!/usr/bin/env python3
"""Expected nodal displacements for the reference's own input decks.

Runs only in the development container: reads the reference's committed
benchmark output (benchmark_XML/level.0/enhanced.strain/benchmark/
Q1P0.3D.b.io0.run, written by the reference's OutputBaseT) and keeps the
D_X, D_Y, D_Z columns of every print step as tests/golden/q1p0_3d_b_run.npz
(time, u[step][node][3]).  The deck and its geometry are the reference's
input files, copied unchanged to tests/golden/decks/.
"""
import os
import re
import sys

import numpy as np

REF = "/root/reference/benchmark_XML/level.0/enhanced.strain/benchmark/Q1P0.3D.b.io0.run"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_run(text, ncols=3):
    """[(time, u[nnd][ncols])] from a .run file's nodal data blocks."""
    out = []
    lines = text.splitlines()
    t = None
    i = 0
    while i < len(lines):
        m = re.match(r"^ Time\. .*= (\S+)", lines[i])
        if m:
            t = float(m.group(1))
        m = re.match(r"^ Number of nodal points\. .*= (\d+)", lines[i])
        if m and t is not None:
            n = int(m.group(1))
            j = i + 1
            while not lines[j].strip().startswith("index"):
                j += 1
            rows = [lines[j + 1 + k].split() for k in range(n)]
            out.append((t, np.array([[float(v) for v in r[2:2 + ncols]] for r in rows])))
            i = j + n
        i += 1
    return out


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else REF
    steps = parse_run(open(src).read())
    np.savez(os.path.join(HERE, "q1p0_3d_b_run.npz"), time=np.array([s[0] for s in steps]),
             u=np.array([s[1] for s in steps]))
    print(f"{len(steps)} steps, {steps[0][1].shape[0]} nodes")


if __name__ == "__main__":
    main()
