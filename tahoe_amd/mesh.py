"""Seeded structured hexahedral meshes and boundary conditions.

The node/element numbering follows the TahoeII cube inputs used by the
reference's benchmark decks (e.g. ``benchmark_XML/level.0/enhanced.strain/
geometry/cube.0.geom``): nodes run x-fastest, then y, then z; element nodes
are the bottom face counter-clockwise followed by the top face, matching the
hexahedron vertex order ``ra/sa/ta`` of ``toolbox/src/geometry/HexahedronT.cpp:23-25``.

Random numbers come from a counter-based SplitMix64 (Steele, Lea & Flood,
"Fast splittable pseudorandom number generators", OOPSLA 2014): draw ``k`` of
stream ``seed`` is the SplitMix64 output for state ``seed + (k+1)*golden``.
Because it is counter based the same mesh can be regenerated on any host and
vectorised with numpy.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, counters: np.ndarray) -> np.ndarray:
    """SplitMix64 output for each counter value (uint64 array)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + (counters.astype(np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, counters: np.ndarray) -> np.ndarray:
    """Doubles in [0, 1) with 53 random bits."""
    return (splitmix64(seed, counters) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


class HexMesh:
    """A hexahedral mesh plus the node sets the boundary conditions refer to.

    Attributes
    ----------
    X : (nnd, 3) float64  reference coordinates
    conn : (nel, 8) int32  0-based connectivity
    mat : (nel,) int32  material (element block) index, 0-based
    nodesets : dict[int, ndarray]  node-set ID -> 0-based node numbers
    shape : (nx, ny, nz) element counts of the structured grid
    """

    def __init__(self, X, conn, mat, nodesets, shape):
        self.X = np.ascontiguousarray(X, dtype=np.float64)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32)
        self.mat = np.ascontiguousarray(mat, dtype=np.int32)
        self.nodesets = nodesets
        self.shape = shape

    @property
    def nnd(self) -> int:
        return self.X.shape[0]

    @property
    def nel(self) -> int:
        return self.conn.shape[0]


def cube(n: int, seed: int = 2024, perturb: float = 0.05, length: float = 1.0,
         nxyz=None) -> HexMesh:
    """Unit cube of n^3 trilinear hexes (or nx*ny*nz when ``nxyz`` is given).

    Interior nodes (not on any face) are moved by U(-perturb*h, +perturb*h)
    per coordinate, h the element edge.  Draw index of coordinate c of node a
    is 3a+c.  Node set 1 is the z=0 face, node set 2 the z=L face.
    """
    nx, ny, nz = (n, n, n) if nxyz is None else nxyz
    hx, hy, hz = length / nx, length / ny, length / nz
    i = np.arange(nx + 1)
    j = np.arange(ny + 1)
    k = np.arange(nz + 1)
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    I = I.ravel(); J = J.ravel(); K = K.ravel()
    X = np.stack([I * hx, J * hy, K * hz], axis=1).astype(np.float64)
    nnd = X.shape[0]
    if perturb != 0.0:
        interior = (I > 0) & (I < nx) & (J > 0) & (J < ny) & (K > 0) & (K < nz)
        ids = np.nonzero(interior)[0].astype(np.uint64)
        h = np.array([hx, hy, hz])
        for c in range(3):
            u = uniform01(seed, np.uint64(3) * ids + np.uint64(c))
            X[ids.astype(np.int64), c] += (2.0 * u - 1.0) * perturb * h[c]
    # element connectivity
    ei = np.arange(nx); ej = np.arange(ny); ek = np.arange(nz)
    EK, EJ, EI = np.meshgrid(ek, ej, ei, indexing="ij")
    EI = EI.ravel(); EJ = EJ.ravel(); EK = EK.ravel()

    def nid(a, b, c):
        return a + (nx + 1) * (b + (ny + 1) * c)

    conn = np.stack([
        nid(EI, EJ, EK), nid(EI + 1, EJ, EK), nid(EI + 1, EJ + 1, EK), nid(EI, EJ + 1, EK),
        nid(EI, EJ, EK + 1), nid(EI + 1, EJ, EK + 1), nid(EI + 1, EJ + 1, EK + 1), nid(EI, EJ + 1, EK + 1),
    ], axis=1).astype(np.int32)
    mat = np.zeros(conn.shape[0], dtype=np.int32)
    nodesets = {
        1: np.nonzero(K == 0)[0].astype(np.int32),
        2: np.nonzero(K == nz)[0].astype(np.int32),
    }
    del nnd
    return HexMesh(X, conn, mat, nodesets, (nx, ny, nz))


def boundary_nodes(mesh: HexMesh) -> np.ndarray:
    """0-based numbers of the nodes on the outer faces of a structured cube."""
    nx, ny, nz = mesh.shape
    a = np.arange(mesh.nnd)
    I = a % (nx + 1)
    J = (a // (nx + 1)) % (ny + 1)
    K = a // ((nx + 1) * (ny + 1))
    on = (I == 0) | (I == nx) | (J == 0) | (J == ny) | (K == 0) | (K == nz)
    return np.nonzero(on)[0].astype(np.int32)


def rve(n: int, seed: int, p_phase2: float = 0.3, hmax: float = 0.02):
    """Voxel two-phase RVE: n^3 unit-cube hexes (unperturbed), phase 2 per
    element ~ Bernoulli(p_phase2), macro displacement gradient H with entries
    U(-hmax, hmax).  Draw indices: element e phase -> e; H entry m -> 2^40 + m.

    Returns (mesh, H) with mesh.mat in {0, 1}.
    """
    mesh = cube(n, seed=seed, perturb=0.0)
    ph = uniform01(seed, np.arange(mesh.nel, dtype=np.uint64)) < p_phase2
    mesh.mat = ph.astype(np.int32)
    hu = uniform01(seed, np.uint64(1 << 40) + np.arange(9, dtype=np.uint64))
    H = ((2.0 * hu - 1.0) * hmax).reshape(3, 3)
    return mesh, H


class KBC:
    """Kinematic boundary conditions in the reference's card form
    (``tahoe/src/primitives/KBC_CardT.cpp``): per (node, dof) a code
    (fixed or prescribed displacement) and a value scaled by a schedule.

    ``nodes``/``dofs`` are 0-based, ``values`` are the card values; the
    prescribed displacement at time t is ``value * schedule(t)``.  With the
    piecewise-linear schedule (0,0)-(1,1) used throughout, schedule(t) = t.
    Fixed cards have value 0.
    """

    def __init__(self, nodes, dofs, values):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        self.dofs = np.ascontiguousarray(dofs, dtype=np.int32)
        self.values = np.ascontiguousarray(values, dtype=np.float64)

    def __len__(self):
        return len(self.nodes)


def kbc_uniaxial(mesh: HexMesh, stretch: float) -> KBC:
    """z=0 face fixed in x, y, z; z=L face u_z = stretch * schedule."""
    b = mesh.nodesets[1]
    t = mesh.nodesets[2]
    nodes = np.concatenate([b, b, b, t])
    dofs = np.concatenate([np.zeros_like(b), np.ones_like(b), np.full_like(b, 2), np.full_like(t, 2)])
    vals = np.concatenate([np.zeros(3 * len(b)), np.full(len(t), stretch)])
    return KBC(nodes, dofs, vals)


def kbc_affine(mesh: HexMesh, H: np.ndarray) -> KBC:
    """All outer-face nodes prescribed u = H X * schedule (all three dofs)."""
    bn = boundary_nodes(mesh)
    U = mesh.X[bn] @ H.T
    nodes = np.repeat(bn, 3)
    dofs = np.tile(np.arange(3, dtype=np.int32), len(bn))
    return KBC(nodes, dofs, U.ravel())


def equation_numbers(nnd: int, kbc: KBC) -> tuple[np.ndarray, int]:
    """Node-major, dof-minor, 1-based numbering of the free dofs; dofs with a
    kinematic card get -1 (``NodeManagerT::SetEquationNumbers``,
    ``tahoe/src/nodes/NodeManagerT.cpp:712-760``; ``FieldT::InitEquations``,
    ``tahoe/src/nodes/FieldT.cpp:635-658``)."""
    eq = np.zeros((nnd, 3), dtype=np.int32)
    eq[kbc.nodes, kbc.dofs] = -1
    free = eq.ravel() == 0
    flat = eq.ravel()
    flat[free] = np.arange(1, int(free.sum()) + 1, dtype=np.int32)
    return flat.reshape(nnd, 3), int(free.sum())
