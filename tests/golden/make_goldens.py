#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ from the reference's own CPU build.

Runs only in the development container, where the reference (Tahoe, read-only
at /root/reference) was compiled from its own sources out of tree; the binary
is passed with --tahoe (default /tmp/tahoe_ref_build/bin/tahoe).  Nothing here
runs on the GPU host; the tests read the committed .npz files only.

For each case a seeded mesh from tahoe_amd.mesh is written as a TahoeII .geom
file (coordinates with 17 significant digits, so the reference reads exactly
the same doubles), an XML deck selects the element/material/solver, and the
reference is run twice:

  check_code="print_RHS"  -> FEManagerT::WriteSystemConfig at DBL_DIG after
                             every FormRHS (tahoe/src/main/FEManagerT.cpp:415,1213):
                             nodal u and residual at every Newton iterate.
  check_code="print_LHS"  -> MSRMatrixT::PrintLHS (MSRMatrixT.cpp:449-471):
                             the assembled upper-triangle MSR matrix as RCV
                             (1-based, 15 significant digits, exact zeros
                             dropped) before every linear solve.

Fixture arrays (all float64/int32, numpy .npz without pickles):
  X, conn, mat, kbc_nodes, kbc_dofs, kbc_values, E, nu, num_steps
  eqnos                 (nnd,3)  reference equation numbers (print_eqnos)
  it_u, it_f            (m,nnd,3) u and residual at every iterate of step 1
  lhs_ptr, lhs_r, lhs_c, lhs_v   concatenated RCV dumps of step-1 solves
  step_u                (num_steps,nnd,3) converged u per step (16 digits)
  step_iters            (num_steps,) Newton corrections per step
"""
from __future__ import annotations

import argparse
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tahoe_amd import mesh as M  # noqa: E402

SCHED = ('<time num_steps="{ns}" output_inc="1" time_step="{dt!r}"><schedule_function>'
         '<piecewise_linear><OrderedPair x="0.0" y="0.0"/><OrderedPair x="1.0" y="1.0"/>'
         '</piecewise_linear></schedule_function></time>')


def write_geom(path, mesh: M.HexMesh, nodesets: dict):
    """TahoeII text geometry; one element set per material (block IDs 1..)."""
    blocks = sorted(set(mesh.mat.tolist()))
    with open(path, "w") as f:
        f.write("*version\n1.0\n*title\ngolden\n*dimensions\n")
        f.write(f"{mesh.nnd}\n3\n{len(blocks)}\n")
        for b in blocks:
            f.write(f"{b + 1} {int((mesh.mat == b).sum())} 8\n")
        f.write(f"{len(nodesets)}\n")
        for sid in sorted(nodesets):
            f.write(f"{sid} {len(nodesets[sid])}\n")
        f.write("0\n*nodesets\n")
        for sid in sorted(nodesets):
            ids = nodesets[sid] + 1
            f.write(f"*set\n{len(ids)}\n")
            for i in range(0, len(ids), 10):
                f.write(" ".join(str(v) for v in ids[i:i + 10]) + "\n")
        f.write("*sidesets\n*elements\n")
        for b in blocks:
            els = np.nonzero(mesh.mat == b)[0]
            f.write(f"*set\n{len(els)}\n8\n")
            for k, e in enumerate(els):
                f.write(f"{k + 1} " + " ".join(str(v + 1) for v in mesh.conn[e]) + "\n")
        f.write(f"*nodes\n{mesh.nnd}\n3\n")
        for a in range(mesh.nnd):
            x, y, z = mesh.X[a]
            f.write(f"{a + 1} {x:.17g} {y:.17g} {z:.17g}\n")


def kbc_cards(kbc: M.KBC):
    """Group KBC entries into node sets + cards.  Cards with the same (dof,
    value) share one node set.  Returns (nodesets, xml)."""
    groups = {}
    for n, d, v in zip(kbc.nodes.tolist(), kbc.dofs.tolist(), kbc.values.tolist()):
        groups.setdefault((d, v), []).append(n)
    nodesets = {}
    cards = []
    for sid, ((d, v), nodes) in enumerate(sorted(groups.items()), start=1):
        nodesets[sid] = np.array(sorted(set(nodes)), dtype=np.int32)
        if v == 0.0:
            cards.append(f'<kinematic_BC dof="{d + 1}" node_ID="{sid}"/>')
        else:
            cards.append(f'<kinematic_BC dof="{d + 1}" node_ID="{sid}" schedule="1" type="u" value="{v!r}"/>')
    return nodesets, "".join(cards)


def write_xml(path, geom, element, E, nu, ns, check, kbc_xml, nblocks):
    mats = "".join(
        f'<large_strain_element_block><block_ID_list><String value="{b + 1}"/></block_ID_list>'
        f'<large_strain_material_3D><Simo_isotropic density="1.0"><E_and_nu Poisson_ratio="{nu[b]!r}" '
        f'Young_modulus="{E[b]!r}"/></Simo_isotropic></large_strain_material_3D></large_strain_element_block>'
        for b in range(nblocks))
    xml = ('<?xml version="1.0"?>\n'
           f'<tahoe geometry_file="{geom}">' + SCHED.format(ns=ns, dt=1.0 / ns) +
           '<nodes><field field_name="displacement"><dof_labels><String value="D_X"/><String value="D_Y"/>'
           '<String value="D_Z"/></dof_labels>' + kbc_xml + '</field></nodes>'
           f'<element_list><{element} field_name="displacement"><hexahedron num_ip="8"/>'
           '<solid_element_nodal_output displacements="1"/>' + mats + f'</{element}></element_list>'
           '<nonlinear_solver abs_tolerance="1.0e-14" divergence_tolerance="10.0" max_iterations="20" '
           f'rel_tolerance="1.0e-10" check_code="{check}" print_eqnos="true"><SPOOLES_matrix/></nonlinear_solver>'
           '</tahoe>\n')
    with open(path, "w") as f:
        f.write(xml)


def run(tahoe, wd, xml):
    r = subprocess.run([tahoe, "-f", xml], cwd=wd, capture_output=True, text=True, timeout=3600)
    if r.returncode != 0 or "End Execution" not in r.stdout:
        raise RuntimeError(f"reference run failed in {wd}:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    return r.stdout


def parse_eqnos(out_text, nnd):
    i = out_text.index("N o d a l   E q u a t i o n   N u m b e r s")
    lines = out_text[i:].splitlines()
    eq = np.zeros((nnd, 3), dtype=np.int32)
    seen = 0
    for ln in lines:
        p = ln.split()
        if len(p) == 5 and all(re.fullmatch(r"-?\d+", t) for t in p):
            eq[int(p[0]) - 1] = [int(p[2]), int(p[3]), int(p[4])]
            seen += 1
            if seen == nnd:
                break
    assert seen == nnd
    return eq


def parse_config_blocks(out_text, nnd):
    """[(time, u(nnd,3), f(nnd,3))] in print order."""
    blocks = []
    lines = out_text.splitlines()
    i = 0
    while i < len(lines):
        m = re.match(r"^\s*time = (\S+)\s*$", lines[i])
        if m and i + 2 < len(lines) and lines[i + 1].strip().startswith("nodes ="):
            t = float(m.group(1))
            rows = lines[i + 3:i + 3 + nnd]
            a = np.array([[float(v) for v in r.split()] for r in rows])
            blocks.append((t, a[:, 8:11].copy(), a[:, 11:14].copy()))
            i += 3 + nnd
        else:
            i += 1
    return blocks


def parse_lhs(path):
    a = np.loadtxt(path, ndmin=2)
    return a[:, 0].astype(np.int32), a[:, 1].astype(np.int32), a[:, 2].copy()


def make_case(tahoe, name, mesh, kbc, element, E, nu, ns, outdir, want_lhs=True):
    nodesets, kbc_xml = kbc_cards(kbc)
    nb = int(mesh.mat.max()) + 1
    res = {}
    with tempfile.TemporaryDirectory(prefix=f"gold_{name}_") as wd:
        write_geom(os.path.join(wd, "mesh.geom"), mesh, nodesets)
        write_xml(os.path.join(wd, "rhs.xml"), "mesh.geom", element, E, nu, ns, "print_RHS", kbc_xml, nb)
        run(tahoe, wd, "rhs.xml")
        out = open(os.path.join(wd, "rhs.out")).read()
        eq = parse_eqnos(out, mesh.nnd)
        blocks = parse_config_blocks(out, mesh.nnd)
        times = [round(b[0] * ns) for b in blocks]
        step_u, step_iters = [], []
        for s in range(1, ns + 1):
            idx = [k for k, t in enumerate(times) if t == s]
            step_u.append(blocks[idx[-1]][1])
            step_iters.append(len(idx) - 1)
        idx1 = [k for k, t in enumerate(times) if t == 1]
        res.update(eqnos=eq, step_u=np.array(step_u), step_iters=np.array(step_iters, dtype=np.int32),
                   it_u=np.array([blocks[k][1] for k in idx1]), it_f=np.array([blocks[k][2] for k in idx1]))
        if want_lhs:
            write_xml(os.path.join(wd, "lhs.xml"), "mesh.geom", element, E, nu, ns, "print_LHS", kbc_xml, nb)
            run(tahoe, wd, "lhs.xml")
            nsolve = step_iters[0]
            rs, cs, vs, ptr = [], [], [], [0]
            for k in range(nsolve):
                r, c, v = parse_lhs(os.path.join(wd, f"MSRMatrixT.LHS.{k}"))
                rs.append(r); cs.append(c); vs.append(v); ptr.append(ptr[-1] + len(r))
            res.update(lhs_ptr=np.array(ptr, dtype=np.int64), lhs_r=np.concatenate(rs),
                       lhs_c=np.concatenate(cs), lhs_v=np.concatenate(vs))
    np.savez_compressed(
        os.path.join(outdir, f"{name}.npz"), X=mesh.X, conn=mesh.conn, mat=mesh.mat,
        kbc_nodes=kbc.nodes, kbc_dofs=kbc.dofs, kbc_values=kbc.values,
        E=np.array(E, dtype=np.float64), nu=np.array(nu, dtype=np.float64), num_steps=np.int32(ns),
        element=np.array(element), **res)
    print(f"{name}: nnd={mesh.nnd} nel={mesh.nel} n_eq={int(eq.max())} iters={step_iters}", flush=True)


CASES = {
    # name: (builder, element, E, nu, num_steps, want_lhs)
    "tl_n2": (lambda: (M.cube(2, seed=11), 0.05), "total_lagrangian", [1000.0], [0.3], 2, True),
    "tl_n4": (lambda: (M.cube(4, seed=12), 0.05), "total_lagrangian", [1000.0], [0.3], 5, True),
    "tl_n8": (lambda: (M.cube(8, seed=13), 0.05), "total_lagrangian", [1000.0], [0.3], 5, False),
    "tl_n16": (lambda: (M.cube(16, seed=2024), 0.05), "total_lagrangian", [1000.0], [0.3], 5, False),
    "tl_n4_s20": (lambda: (M.cube(4, seed=14), 0.20), "total_lagrangian", [1000.0], [0.3], 10, False),
    "q1p0_n2": (lambda: (M.cube(2, seed=21), 0.05), "updated_lagrangian_Q1P0", [1000.0], [0.3], 2, True),
    "q1p0_n4": (lambda: (M.cube(4, seed=22), 0.05), "updated_lagrangian_Q1P0", [1000.0], [0.3], 2, True),
    "q1p0_n4_inc": (lambda: (M.cube(4, seed=23), 0.05), "updated_lagrangian_Q1P0", [1000.0], [0.499], 10, True),
}


def build_case(name):
    if name.startswith("rve"):
        n = int(name.split("_n")[1])
        mesh, H = M.rve(n, seed=100 + n)
        return mesh, M.kbc_affine(mesh, H)
    (mk, el, E, nu, ns, lhs) = CASES[name]
    mesh, stretch = mk()
    return mesh, M.kbc_uniaxial(mesh, stretch)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--tahoe", default="/tmp/tahoe_ref_build/bin/tahoe")
    ap.add_argument("--out", default=HERE)
    ap.add_argument("cases", nargs="*")
    a = ap.parse_args()
    names = a.cases or list(CASES) + ["rve_n3", "rve_n4"]
    for name in names:
        mesh, kbc = build_case(name)
        if name.startswith("rve"):
            make_case(a.tahoe, name, mesh, kbc, "total_lagrangian", [1000.0, 10000.0], [0.3, 0.3], 2,
                      a.out, want_lhs=True)
        else:
            (_, el, E, nu, ns, lhs) = CASES[name]
            make_case(a.tahoe, name, mesh, kbc, el, E, nu, ns, a.out, want_lhs=lhs)


if __name__ == "__main__":
    main()
