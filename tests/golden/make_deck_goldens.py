#!/usr/bin/env python3
"""Expected nodal displacements for the reference's own input decks.

Runs only in the development container: reads the reference's committed
benchmark output (.run files under benchmark_XML/level.0/*/benchmark/,
written by the reference's OutputBaseT) and keeps the D_X, D_Y, D_Z columns
of every print step as tests/golden/<name>_run.npz (time, u[step][node][3]):
  Q1P0.3D.b          enhanced.strain (updated_lagrangian_Q1P0, Simo_isotropic)
  beam.Newton        3D.elastostatic (updated_lagrangian, large_strain_StVenant,
                     force_BC, nonlinear_solver)
  beam.Newton.TotLag 3D.elastostatic (total_lagrangian, otherwise the same)
  beam.PCG           3D.elastostatic (updated_lagrangian, PCG_solver)
The decks and their geometry are the reference's input files, copied
unchanged to tests/golden/decks/ and tests/golden/geometry/.

It also keeps, as tests/golden/deck_equations.json, the "Total number of
equations" that the reference's .out files print for 3D hexahedron decks,
with each deck's geometry file and kinematic_BC cards (node set, dof).
(The same files print the size of the profile matrix, but the profile
matrix renumbers the equations first -- CCSMatrixT::RenumberEquations,
FEManagerT.cpp:2266 -- so those sizes say nothing about the MSR numbering,
which keeps the original order, MSRMatrixT.h:71.)
"""
import json
import os
import re
import sys

import numpy as np

LEVEL0 = "/root/reference/benchmark_XML/level.0"
RUNS = {
    "q1p0_3d_a": "enhanced.strain/benchmark/Q1P0.3D.a.io0.run",
    "q1p0_3d_b": "enhanced.strain/benchmark/Q1P0.3D.b.io0.run",
    "beam_newton": "3D.elastostatic/benchmark/beam.Newton.io0.run",
    "beam_newton_totlag": "3D.elastostatic/benchmark/beam.Newton.TotLag.io0.run",
    "beam_pcg": "3D.elastostatic/benchmark/beam.PCG.io0.run",
}
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


BENCH = "/root/reference/benchmark_XML"
# deck -> (deck, .out, geometry fixture)
DECKS = {
    "mat.5.a": ("level.1/material.solid/3D/material.05/mat.5.a.xml",
                "level.1/material.solid/3D/material.05/benchmark/mat.5.a.out", "cube.1.geom"),
    "traction.a": ("level.0/3D.elastostatic/traction.a.xml", "level.0/3D.elastostatic/benchmark/traction.a.out",
                   "cube.1.geom"),
    "beam.Newton.TotLag": ("level.0/3D.elastostatic/beam.Newton.TotLag.xml",
                           "level.0/3D.elastostatic/benchmark/beam.Newton.TotLag.out", "beam.1.geom"),
    "Q1P0.3D.a": ("level.0/enhanced.strain/Q1P0.3D.a.xml", "level.0/enhanced.strain/benchmark/Q1P0.3D.a.out",
                  "beam.1.geom"),
}


def deck_equations(root=BENCH):
    out = {}
    for name, (xml, dot_out, geom) in DECKS.items():
        text = open(os.path.join(root, dot_out)).read()
        m = re.search(r"^ Total number of equations[ .]*= (\d+)", text, flags=re.M)
        deck = open(os.path.join(root, xml)).read()
        out[name] = {"deck": xml, "out": dot_out, "geometry": geom, "n_eq": int(m.group(1)),
                     "kbc": [[int(n), int(d)] for d, n in
                             re.findall(r'<kinematic_BC dof="(\d)" node_ID="(\d+)"', deck)]}
    return out


def main():
    with open(os.path.join(HERE, "deck_equations.json"), "w") as f:
        json.dump(deck_equations(), f, indent=1)
    root = sys.argv[1] if len(sys.argv) > 1 else LEVEL0
    for name, rel in RUNS.items():
        steps = parse_run(open(os.path.join(root, rel)).read())
        np.savez(os.path.join(HERE, name + "_run.npz"), time=np.array([s[0] for s in steps]),
                 u=np.array([s[1] for s in steps]))
        print(f"{name}: {len(steps)} steps, {steps[0][1].shape[0]} nodes")


if __name__ == "__main__":
    main()
