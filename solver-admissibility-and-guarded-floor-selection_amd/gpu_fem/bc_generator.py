"""Boundary conditions and load vector of a ProblemSpec (bc_generator.py:63-447
of the reference, 3D branch): supports and loads snap to the nearest mesh node,
faces enumerate their nodes, distributed loads use trapezoidal tributary
areas, loads on fixed DOFs are zeroed."""

from __future__ import annotations

import numpy as np

from .problem_spec import (DistributedLoad, EdgeNodesLoad, EdgeSupport, PointLoad, PointSupport,
                           ProblemSpec)


class BCResult:
    __slots__ = ("fixed_dofs", "free_dofs", "F", "ndof", "node_coords", "passive_mask")

    def __init__(self, fixed_dofs, free_dofs, F, ndof, node_coords):
        self.fixed_dofs = fixed_dofs
        self.free_dofs = free_dofs
        self.F = F
        self.ndof = ndof
        self.node_coords = node_coords
        self.passive_mask = None


def _coords(spec: ProblemSpec) -> np.ndarray:
    n = np.arange((spec.nelx + 1) * (spec.nely + 1) * (spec.nelz + 1))
    iz = n % (spec.nelz + 1)
    iy = (n // (spec.nelz + 1)) % (spec.nely + 1)
    ix = n // ((spec.nelz + 1) * (spec.nely + 1))
    return np.column_stack([ix * (spec.Lx / spec.nelx), iy * (spec.Ly / spec.nely),
                            iz * (spec.Lz / spec.nelz)])


def _node(i, j, k, spec):
    return (i * (spec.nely + 1) + j) * (spec.nelz + 1) + k


def _face(face: str, spec: ProblemSpec) -> np.ndarray:
    X, Y, Z = spec.nelx, spec.nely, spec.nelz
    a = {"left": [(0, j, k) for j in range(Y + 1) for k in range(Z + 1)],
         "right": [(X, j, k) for j in range(Y + 1) for k in range(Z + 1)],
         "bottom": [(i, 0, k) for i in range(X + 1) for k in range(Z + 1)],
         "top": [(i, Y, k) for i in range(X + 1) for k in range(Z + 1)],
         "front": [(i, j, 0) for i in range(X + 1) for j in range(Y + 1)],
         "back": [(i, j, Z) for i in range(X + 1) for j in range(Y + 1)]}
    if face not in a:
        raise ValueError(f"Unknown 3D face: {face}")
    return np.array([_node(i, j, k, spec) for i, j, k in a[face]], dtype=np.int64)


def _nearest(coords, p) -> int:
    return int(np.argmin(np.linalg.norm(coords - np.asarray(p, dtype=float)[None, :], axis=1)))


def _tributary(c: np.ndarray, spec: ProblemSpec, face: str) -> np.ndarray:
    if face in ("left", "right"):
        h1, h2, c1, c2 = spec.Ly / spec.nely, spec.Lz / spec.nelz, c[:, 1], c[:, 2]
    elif face in ("bottom", "top"):
        h1, h2, c1, c2 = spec.Lx / spec.nelx, spec.Lz / spec.nelz, c[:, 0], c[:, 2]
    else:
        h1, h2, c1, c2 = spec.Lx / spec.nelx, spec.Ly / spec.nely, c[:, 0], c[:, 1]

    def one(cv, h):
        w = np.full(len(cv), h)
        w[np.abs(cv - cv.min()) < h * 0.01] = h / 2
        w[np.abs(cv - cv.max()) < h * 0.01] = h / 2
        return w

    return one(c1, h1) * one(c2, h2)


def generate_bc(spec: ProblemSpec) -> BCResult:
    errors = spec.validate()
    if errors:
        raise ValueError("ProblemSpec validation failed:\n" + "\n".join(f"  - {e}" for e in errors))
    coords = _coords(spec)
    ndof = 3 * len(coords)
    fixed = set()
    for s in spec.supports:
        nodes = _face(s.edge, spec) if isinstance(s, EdgeSupport) else [_nearest(coords, (s.x, s.y, s.z))]
        for n in nodes:
            fixed.update(3 * int(n) + o for o in s.dof_mask(3))
    fixed = np.array(sorted(fixed), dtype=np.int64)
    free = np.setdiff1d(np.arange(ndof), fixed)
    F = np.zeros(ndof)
    for ld in spec.loads:
        if isinstance(ld, PointLoad):
            n = _nearest(coords, (ld.x, ld.y, ld.z))
            F[3 * n:3 * n + 3] += (ld.fx, ld.fy, ld.fz)
        elif isinstance(ld, DistributedLoad):
            nodes = _face(ld.edge, spec)
            w = _tributary(coords[nodes], spec, ld.edge)
            off, sign = {"left": (0, -1.0), "right": (0, 1.0), "bottom": (1, -1.0),
                         "top": (1, 1.0), "front": (2, -1.0), "back": (2, 1.0)}[ld.edge]
            for i, n in enumerate(nodes):
                F[3 * n + off] += ld.magnitude * sign * w[i]
        elif isinstance(ld, EdgeNodesLoad):
            nodes = [_node(spec.nelx, 0, k, spec) for k in range(spec.nelz + 1)]
            for n in nodes:
                F[3 * n + ld.component] += ld.total / len(nodes)
        else:
            raise ValueError(f"unsupported load {ld!r}")
    F[fixed] = 0.0
    if np.max(np.abs(F[free])) < 1e-30:
        raise ValueError("Force vector is zero on all free DOFs. "
                         "Check that loads are not placed on fully-fixed supports.")
    return BCResult(fixed, free, F, ndof, coords)
