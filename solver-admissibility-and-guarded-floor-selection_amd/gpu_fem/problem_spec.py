"""Problem specification records (the subset of problem_spec.py:31-227 that
the solver path needs: geometry, supports, loads)."""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

_DOF_MASKS = {
    "fixed": [0, 1, 2], "pin_x": [0], "pin_y": [1], "pin_z": [2],
    "roller_x": [1, 2], "roller_y": [0, 2], "roller_z": [0, 1],
}


@dataclass
class PointSupport:
    x: float
    y: float
    z: float = 0.0
    constraint: str = "fixed"

    def dof_mask(self, ndim: int = 3) -> list:
        return list(_DOF_MASKS.get(self.constraint, [0, 1, 2]))


@dataclass
class EdgeSupport:
    edge: str
    constraint: str = "fixed"

    def dof_mask(self, ndim: int = 3) -> list:
        return list(_DOF_MASKS.get(self.constraint, [0, 1, 2]))


@dataclass
class PointLoad:
    x: float
    y: float
    z: float = 0.0
    fx: float = 0.0
    fy: float = 0.0
    fz: float = 0.0

    @property
    def magnitude(self) -> float:
        return math.sqrt(self.fx ** 2 + self.fy ** 2 + self.fz ** 2)


@dataclass
class DistributedLoad:
    edge: str
    magnitude: float = -1.0


@dataclass
class EdgeNodesLoad:
    """Unit total force split equally over the nodes of the bottom edge
    (y = 0) of the free end (x = Lx), all z -- the cantilever load of the
    batched/precision workloads.  Not in the reference."""
    total: float = -1.0
    component: int = 1


@dataclass
class ProblemSpec:
    Lx: float = 2.0
    Ly: float = 1.0
    Lz: float = 0.0
    nelx: int = 60
    nely: int = 30
    nelz: int = 0
    E: float = 1.0
    nu: float = 0.3
    volfrac: float = 0.5
    supports: list = field(default_factory=list)
    loads: list = field(default_factory=list)
    passive_regions: list = field(default_factory=list)
    max_iter: Optional[int] = None
    rmin: Optional[float] = None

    @property
    def is_3d(self) -> bool:
        return self.Lz > 0 and self.nelz > 0

    @property
    def ndim(self) -> int:
        return 3 if self.is_3d else 2

    def validate(self) -> list:
        errors = []
        if self.Lx <= 0 or self.Ly <= 0:
            errors.append("Domain dimensions Lx, Ly must be positive.")
        if not self.is_3d:
            errors.append("Only 3D hex problems are supported by this solver path.")
        if self.nelx < 1 or self.nely < 1 or self.nelz < 1:
            errors.append("Mesh must have at least 1 element per axis.")
        if not self.supports:
            errors.append("At least one support (fixed DOF) is required.")
        if not self.loads:
            errors.append("At least one load is required.")
        return errors
