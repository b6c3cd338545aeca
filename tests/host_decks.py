"""Helpers shared by the host-layer tests: write a golden fixture as a Tahoe
XML deck + TahoeII geometry (the same writer that produced the fixture's
reference run) and run or parse the tahoe_hip driver."""
import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "golden"))
EXE = os.path.join(ROOT, "tahoe_amd", "bin", "tahoe_hip")
DECKS = os.path.join(HERE, "golden", "decks")


def write_fixture_deck(g, folder, name="deck", element=None):
    import make_goldens as MG
    from tahoe_amd import mesh as M
    m = M.HexMesh(g["X"], g["conn"], g["mat"], {}, None)
    k = M.KBC(g["kbc_nodes"], g["kbc_dofs"], g["kbc_values"])
    ns, kx = MG.kbc_cards(k)
    MG.write_geom(os.path.join(folder, name + ".geom"), m, ns)
    xml = os.path.join(folder, name + ".xml")
    MG.write_xml(xml, name + ".geom", element or str(g["element"]), [float(v) for v in g["E"]],
                 [float(v) for v in g["nu"]], int(g["num_steps"]), "no_check", kx, int(m.mat.max()) + 1)
    return xml


def run(args, timeout=120):
    return subprocess.run([EXE] + list(args), capture_output=True, text=True, timeout=timeout)


def parse_run(path):
    """[(time, u[nnd][3])] from a .run file (the reference's layout)."""
    out, t = [], None
    lines = open(path).read().splitlines()
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
            out.append((t, np.array([[float(v) for v in lines[j + 1 + k].split()[2:5]] for k in range(n)])))
            i = j + n
        i += 1
    return out


def read_geom(path):
    """Nodes [nnd][3], hexahedra [nel][8] (0-based) and node sets {ID: 0-based
    nodes} of a TahoeII text geometry file (the sections *nodesets,
    *elements, *nodes; '#' starts a comment).  A block whose first token is a
    file name instead of a count is read from that file, beside the
    geometry file (TahoeII's external blocks, e.g. cube.1.elem)."""
    folder = os.path.dirname(path)

    def tokens(p):
        out = []
        for line in open(p):
            out += line.split("#")[0].split()
        return out

    toks = []
    for t in tokens(path):
        if not t.startswith("*") and not t.lstrip("+-").replace(".", "", 1).replace("e", "", 1).replace(
                "-", "").replace("+", "").isdigit() and os.path.exists(os.path.join(folder, t)):
            toks += tokens(os.path.join(folder, t))
        else:
            toks.append(t)
    i = 0
    sets, X, conn = [], None, []

    def num():
        nonlocal i
        i += 1
        return toks[i - 1]
    while i < len(toks):
        t = num()
        if t == "*nodesets":
            while i < len(toks) and toks[i] == "*set":
                num()
                n = int(num())
                sets.append([int(num()) - 1 for _ in range(n)])
        elif t == "*elements":
            while i < len(toks) and toks[i] == "*set":
                num()
                nel, nen = int(num()), int(num())
                for _ in range(nel):
                    num()
                    conn.append([int(num()) - 1 for _ in range(nen)])
        elif t == "*nodes":
            nnd, nsd = int(num()), int(num())
            X = np.zeros((nnd, nsd))
            for a in range(nnd):
                num()
                X[a] = [float(num()) for _ in range(nsd)]
    return X, np.array(conn, dtype=np.int32), {k + 1: np.array(v, dtype=np.int32) for k, v in enumerate(sets)}
