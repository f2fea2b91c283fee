"""Synthetic single-patch quadratic B-spline meshes in IceFEM's mesh format.

IceFEM meshes are generated in MATLAB with igafem and stored in HDF5
(TestCases/TriAxialCompression/Meshgen/GenerateMesh.m:160-372, SaveToHDF.m:1-83):
control points numbered x-fastest (GenerateMesh.m:191-201), elements x-fastest with
their nodes x-fastest (:286-293), and per-element Bezier extraction operators
C = kron(Cz, kron(Cy, Cx)) (bezierExtraction3D.m:17-27), stored so that
ElemData(j,k) = C(j,k) (mesh/Groups/elementgroup.cpp:62).  This module builds the same
arrays for an axis-aligned box with open uniform knot vectors (C1 quadratic,
nel+2 control points per direction, unit weights) so the engine can run without the
MATLAB toolchain.  The 1D extraction follows Borden, Scott, Evans & Hughes (2011),
"Isogeometric finite element data structures based on Bezier extraction", Algorithm 1.
"""
from dataclasses import dataclass, field

import numpy as np


def open_uniform_knots(nel, p=2):
    """[0]*(p+1) + interior + [1]*(p+1); nel elements, nel+p basis functions."""
    interior = [i / nel for i in range(1, nel)]
    return np.array([0.0] * (p + 1) + interior + [1.0] * (p + 1))


def bezier_extraction_1d(knots, p=2):
    """Element extraction operators C_e ((p+1)x(p+1)), N_e = C_e B (Borden et al. 2011,
    Algorithm 1; rows: B-spline functions of the element, columns: Bernstein)."""
    U = [None] + list(map(float, knots))  # 1-based as in the paper
    m = len(knots)
    a, b, nb = p + 1, p + 2, 1
    Cs = {1: np.eye(p + 1)}
    while b < m:
        Cs[nb + 1] = np.eye(p + 1)
        i = b
        while b < m and U[b + 1] == U[b]:
            b += 1
        mult = b - i + 1
        if mult < p:
            numer = U[b] - U[a]
            alphas = [0.0] * (p + 1)
            for j in range(p, mult, -1):
                alphas[j - mult] = numer / (U[a + j] - U[a])
            r = p - mult
            for j in range(1, r + 1):
                save = r - j + 1
                s = mult + j
                Cn = Cs[nb]
                for k in range(p + 1, s, -1):
                    alpha = alphas[k - s]
                    Cn[:, k - 1] = alpha * Cn[:, k - 1] + (1.0 - alpha) * Cn[:, k - 2]
                if b < m:
                    # C^{nb+1}(save:j+save, save) = C^{nb}(p-j+1:p+1, p+1)   (1-based)
                    Cs[nb + 1][save - 1:j + save, save - 1] = Cn[p - j:p + 1, p]
        nb += 1
        if b < m:
            a = b
            b = b + 1
    nel = len(np.unique(knots)) - 1
    return np.stack([Cs[e + 1] for e in range(nel)])


def greville(knots, p=2):
    n = len(knots) - p - 1
    return np.array([np.mean(knots[i + 1:i + p + 1]) for i in range(n)])


@dataclass
class PatchMesh:
    dim: int
    nel: tuple                      # elements per direction
    lengths: tuple                  # box size per direction
    coords: np.ndarray              # (n_nodes, dim)
    weights: np.ndarray             # (n_nodes,)
    elems: np.ndarray               # (n_elems, 3**dim) int64
    extraction_1d: np.ndarray       # (n_elems, dim, 3, 3)
    ipcount1D: int = 3
    origin: tuple = field(default=(0.0, 0.0, 0.0))

    @property
    def n_nodes(self):
        return self.coords.shape[0]

    @property
    def n_elems(self):
        return self.elems.shape[0]

    @property
    def nodes_per_elem(self):
        return self.elems.shape[1]

    def dense_extraction(self, elements=None):
        """(n_elems, n, n) dense operators C = kron(Cz, kron(Cy, Cx)) (bezierExtraction3D.m:23)."""
        E = self.extraction_1d if elements is None else self.extraction_1d[elements]
        if self.dim == 2:
            return np.stack([np.kron(e[1], e[0]) for e in E])
        return np.stack([np.kron(e[2], np.kron(e[1], e[0])) for e in E])


def make_patch(nel, lengths, origin=None):
    """Axis-aligned box meshed by one open-uniform quadratic B-spline patch."""
    nel = tuple(int(v) for v in nel)
    dim = len(nel)
    origin = tuple(origin) if origin is not None else (0.0,) * dim
    knots = [open_uniform_knots(n) for n in nel]
    C1 = [bezier_extraction_1d(k) for k in knots]
    g = [origin[d] + lengths[d] * greville(knots[d]) for d in range(dim)]
    nn = [n + 2 for n in nel]
    # node id = ix + nx*(iy + ny*iz)  (x-fastest)
    grids = np.meshgrid(*g, indexing="ij")
    coords = np.stack([gr.transpose(tuple(range(dim))[::-1]).reshape(-1) for gr in grids], axis=1)
    weights = np.ones(coords.shape[0])
    # elements x-fastest, local nodes x-fastest
    ec = np.meshgrid(*[np.arange(n) for n in nel], indexing="ij")
    ec = [c.transpose(tuple(range(dim))[::-1]).reshape(-1) for c in ec]  # ex, ey[, ez] per element
    loc = np.meshgrid(*[np.arange(3)] * dim, indexing="ij")
    loc = [c.transpose(tuple(range(dim))[::-1]).reshape(-1) for c in loc]  # ax, ay[, az] per local node
    node = np.zeros((len(ec[0]), 3 ** dim), dtype=np.int64)
    stride = 1
    for d in range(dim):
        node += (ec[d][:, None] + loc[d][None, :]) * stride
        stride *= nn[d]
    ext = np.stack([C1[d][ec[d]] for d in range(dim)], axis=1)  # (E, dim, 3, 3)
    return PatchMesh(dim=dim, nel=nel, lengths=tuple(lengths), coords=coords, weights=weights,
                     elems=node, extraction_1d=np.ascontiguousarray(ext), origin=origin)
