"""Named problem presets used by the floor-policy workloads (presets.py of the
reference: cantilever family 34-40 / 137-170, bridge family 43-55 / 265-278,
bracket family 363-369 / 425-440)."""

from __future__ import annotations

from .problem_spec import (DistributedLoad, EdgeNodesLoad, EdgeSupport, PointLoad, PointSupport,
                           ProblemSpec)


def _cantilever(nelx, nely, nelz, load="tip"):
    loads = ([PointLoad(x=2.0, y=0.5, z=0.25, fy=-1.0)] if load == "tip"
             else [EdgeNodesLoad(total=-1.0)])
    return ProblemSpec(Lx=2.0, Ly=1.0, Lz=0.5, nelx=nelx, nely=nely, nelz=nelz, volfrac=0.3,
                       supports=[EdgeSupport(edge="left", constraint="fixed")], loads=loads)


def _bridge(nelx, nely, nelz, zs):
    sup = [PointSupport(x=0.0, y=0.0, z=z, constraint="fixed") for z in zs]
    sup += [PointSupport(x=3.0, y=0.0, z=z, constraint="roller_x") for z in zs]
    return ProblemSpec(Lx=3.0, Ly=1.0, Lz=0.5, nelx=nelx, nely=nely, nelz=nelz, volfrac=0.3,
                       supports=sup, loads=[DistributedLoad(edge="top", magnitude=-1.0)])


def _bracket(nelx, nely, nelz):
    return ProblemSpec(Lx=1.0, Ly=2.0, Lz=0.5, nelx=nelx, nely=nely, nelz=nelz, volfrac=0.3,
                       supports=[EdgeSupport(edge="top", constraint="fixed")],
                       loads=[PointLoad(x=1.0, y=0.0, z=0.25, fy=-0.7, fx=0.3)])


_PRESETS = {
    "cantilever_3d": lambda: _cantilever(24, 12, 6),
    "cantilever_3d_large": lambda: _cantilever(40, 20, 10),
    "cantilever_gpu_medium": lambda: _cantilever(80, 40, 20),
    "cantilever_gpu_large": lambda: _cantilever(120, 60, 30),
    "cantilever_gpu_xlarge": lambda: _cantilever(160, 80, 40),
    "cantilever_gpu_xxlarge": lambda: _cantilever(200, 100, 50),
    "bridge_3d": lambda: _bridge(30, 10, 5, (0.0, 0.25, 0.5)),
    "bridge_gpu_medium": lambda: _bridge(90, 30, 15, (0.0, 0.25, 0.5)),
    "bridge_gpu_large": lambda: _bridge(150, 50, 25, (0.0, 0.125, 0.25, 0.375, 0.5)),
    "bracket_3d": lambda: _bracket(12, 24, 6),
    "bracket_gpu_medium": lambda: _bracket(30, 60, 15),
    "bracket_gpu_500k": lambda: _bracket(80, 160, 40),
    "bracket_gpu_1M": lambda: _bracket(100, 200, 50),
}


def get_preset(name: str) -> ProblemSpec:
    if name not in _PRESETS:
        raise KeyError(f"Unknown preset '{name}'. Available: {sorted(_PRESETS)}")
    return _PRESETS[name]()


def list_presets() -> list:
    return sorted(_PRESETS)


def cantilever(nelx: int, nely: int, nelz: int, load: str = "tip") -> ProblemSpec:
    """Cantilever geometry of the presets at any resolution.

    load="tip": point load fy=-1 at (2.0, 0.5, 0.25) (the reference's);
    load="edge": unit downward load split over the x=L, y=0 edge nodes."""
    if load not in ("tip", "edge"):
        raise ValueError("load must be 'tip' or 'edge'")
    return _cantilever(nelx, nely, nelz, load)
