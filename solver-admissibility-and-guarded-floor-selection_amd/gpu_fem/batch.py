"""Batched fixed-floor guarded solves over states of uneven grid size.

Each state (grid dims, seed, solid probability) is solved at every floor of a
list, each solve accepted only through the recomputed-residual guard
||F - K(E) x|| / ||F|| <= tol.  States run concurrently: ``workers`` host
threads each drive their own HIP stream and build the components of a
state's grid (fine operator, autotuned launch shape, Galerkin hierarchy with
its hipGraph-recorded preconditioner), solve all floors of that state, and
release them.  The preconditioner precision is selectable: "fp32" runs every
multigrid level (fine smoother and Galerkin stencils) in FP32 while the outer
FGMRES, the guard residual and the compliance stay FP64.

Within one state the hierarchy keeps its cached spectral estimate across the
floors (multigrid_v4.py:1170-1192 session semantics, estimated at the first
floor solved).
"""

from __future__ import annotations

import queue
import threading
import time

import numpy as np

from .policy import CANONICAL_STACK, _build_components, _sync
from .presets import cantilever
from .solver_v2 import nrm2, dot


def sample_batch_states(n_states: int = 64, seed: int = 0, lo=(30, 15, 15), hi=(80, 40, 20),
                        probabilities=(0.2, 0.35, 0.5)):
    """Seeded batch: dims drawn uniformly per axis (integers in [lo, hi]),
    state seed = 1000 + index, solid probability cycling over
    ``probabilities``."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_states):
        dims = tuple(int(rng.integers(a, b + 1)) for a, b in zip(lo, hi))
        out.append({"dims": dims, "seed": 1000 + i,
                    "probability": float(probabilities[i % len(probabilities)])})
    return out


def precision_stack(precision: str, n_levels: int = 4) -> dict:
    """Stack keywords of a preconditioner precision.  "fp32": every
    multigrid level (fine smoother, Galerkin stencils) in FP32, and the
    coarsest level solved by the FP64 Cholesky (explicit inverse) up to 8,192
    free DOFs -- at 1M elements its 6,825 DOFs, which the plain FP64 form
    (the reference's rule, dense_chol_max 5,000) solves by 80-200 fixed
    Jacobi-PCG iterations instead.  "bf16": the same, with the fine level
    smoothed through the bf16 matrix-core operator (multigrid_v4.py:1765-1780).
    The outer FGMRES, the guard residual and the compliance stay FP64 in
    every form."""
    if precision == "fp64":
        return {}
    if precision == "fp32":
        return {"fine_smoother": "fp32", "level_precisions": ["fp32"] * n_levels,
                "dense_chol_max": 8192}
    if precision == "bf16":
        return {"fine_smoother": "bf16", "level_precisions": ["bf16"] + ["fp32"] * (n_levels - 1),
                "dense_chol_max": 8192}
    raise ValueError("precision must be 'fp64', 'fp32' or 'bf16'")


class BatchedFloorSolver:
    """Fixed-floor guarded solves of many states on concurrent streams."""

    def __init__(self, *, workers: int = 8, floors=(1e-12, 1e-3, 1e-2), restart: int = 100,
                 maxiter: int = 300, tol: float = 1e-6, penal: float = 3.0, load: str = "edge",
                 precision: str = "fp32", **stack):
        import torch

        self.workers = int(workers)
        self.floors = tuple(float(f) for f in floors)
        self.restart = int(restart)
        self.maxiter = int(maxiter)
        self.tol = float(tol)
        self.penal = float(penal)
        self.load = load
        self.precision = precision
        st = dict(CANONICAL_STACK)
        st.update(precision_stack(precision, st.get("n_levels", 4)))
        st.update(stack)
        self.stack = st
        self.streams = [torch.cuda.Stream() for _ in range(self.workers)]

    def _solve_state(self, state):
        import torch

        dims = tuple(state["dims"])
        n_e = dims[0] * dims[1] * dims[2]
        solid = np.random.default_rng(state["seed"]).random(n_e) < state["probability"]
        t0 = time.perf_counter()
        spec, bc, free, free_gpu, edof_gpu, F, mf, gmg = _build_components(
            cantilever(*dims, load=self.load), **self.stack)
        _sync()
        build_s = time.perf_counter() - t0
        b_norm = nrm2(F)
        rows = []
        for k, rho_min in enumerate(self.floors):
            t1 = time.perf_counter()
            rho = np.where(solid, 1.0, rho_min)
            E = torch.as_tensor(rho_min + (1.0 - rho_min) * rho ** self.penal, device="cuda")
            gmg.setup(E)
            hist = []
            x, iters, reported = gmg.solve(F, tol=self.tol, maxiter=self.maxiter,
                                           restart=self.restart, history=hist)
            final = nrm2(mf.residual(F, x, E)) / b_norm
            compliance = dot(F, x)
            _sync()
            rows.append({
                "dims": "x".join(str(d) for d in dims), "n_elem": n_e, "n_free": len(free),
                "seed": state["seed"], "solid_probability": state["probability"],
                "rho_min": rho_min, "precision": self.precision, "restart": self.restart,
                "iters": int(iters), "solver_reported_converged": int(reported),
                "true_rel_residual": float(final), "accepted": int(final <= self.tol),
                "compliance": float(compliance), "launch_shape": mf.launch_shape,
                "solve_time_s": time.perf_counter() - t1,
                "build_time_s": build_s if k == 0 else 0.0,
            })
        del gmg, mf
        return rows

    def run(self, states):
        """Rows for every (state, floor) in state order."""
        import torch

        results = [None] * len(states)
        todo = queue.Queue()
        for i in range(len(states)):
            todo.put(i)
        errors = []

        def worker(k):
            while not errors:
                try:
                    i = todo.get_nowait()
                except queue.Empty:
                    return
                try:
                    with torch.cuda.stream(self.streams[k]):
                        results[i] = self._solve_state(states[i])
                except BaseException as ex:  # surfaced in the caller
                    errors.append(ex)
                    return

        threads = [threading.Thread(target=worker, args=(k,), daemon=True)
                   for k in range(min(self.workers, max(1, len(states))))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        return [r for rows in results for r in rows]
