"""In-loop SIMP solve on the matrix-free GMG path, plain and guarded.

Counterparts of
  SolverV4 (matrix-free GMG path)      src/gpu_fem/solver_v4.py:20-282
  GuardedAdaptiveFloorSolverV4         experiments/phase5/run_simp_floor_trajectory.py:65-224

``SolverV4.solve(rho_phys, penal) -> (compliance, dc_phys)`` builds
E = Emin + (E0 - Emin) rho^penal, sets the hierarchy up with the operator
diagonal, solves K(E) u = F with PCG or FGMRES (warm-started from the previous
solution, with one cold retry when a warm start does not converge), records
the recomputed true residual, and returns the compliance F.u and the
sensitivity dc = -penal (E0 - Emin) rho^(penal-1) u_e^T KE u_e from one fused
kernel (gf_elem_sensitivity_f64).  Only the matrix-free GMG path of the
reference's solver family is provided; the guarded subclass runs the in-loop
probe -> (probe-converged keep | classify -> full keep solve -> guard) ->
ladder policy on top of it.
"""

from __future__ import annotations

import time
from typing import Optional

import numpy as np

from . import _lib
from .multigrid_v4 import GalerkinMatFreeGMG, _cupy_fgmres
from .pub_simp_solver import grid_from_edof
from .solver_v2 import MatrixFreeKff, _as_dev, _cupy_pcg, _host, dot, nrm2


class SolverV4:
    """Matrix-free full-Galerkin GMG solver for the SIMP loop.

    Positional arguments follow the reference's GPUFEMSolver
    (fem_gpu.py:165-182): edof, row_idx, col_idx, KE_UNIT, free, F, ndof;
    row_idx / col_idx (CSR assembly indices) are accepted and unused -- this
    path never assembles K."""

    def __init__(self, edof, row_idx, col_idx, KE_UNIT, free, F, ndof: int, E0: float = 1.0,
                 Emin: float = 1e-9, backend: str = "auto", cg_tol: float = 1e-5,
                 cg_maxiter: int = 1000, *, grid_dims=None, enable_warm_start: bool = True,
                 enable_matrix_free: bool = True, enable_matfree_gmg: bool = True,
                 matfree_gmg_levels: int = 4, gmg_fine_smoother: str = "fp32",
                 gmg_fine_degree: int = 2, gmg_outer_solver: str = "auto",
                 gmg_restart: int = 32, gmg_smoother_type: str = "chebyshev",
                 gmg_cheb_lower_frac: float = 1.0 / 30.0, gmg_level_precisions=None,
                 gmg_cycle_type: str = "v", gmg_smooth_iters: int = 2,
                 enable_fused_cuda: bool = False, enable_profiling: bool = False, **unused):
        import torch

        if not (enable_matfree_gmg and enable_matrix_free):
            raise ValueError("this SolverV4 provides the matrix-free GMG path only "
                             "(enable_matrix_free=True, enable_matfree_gmg=True)")
        if gmg_fine_smoother not in {"fp64", "fp32", "bf16"}:
            raise ValueError("gmg_fine_smoother must be one of 'fp64', 'fp32', or 'bf16'; "
                             f"got {gmg_fine_smoother!r}.")
        if gmg_outer_solver not in {"auto", "pcg", "fgmres"}:
            raise ValueError("gmg_outer_solver must be 'auto', 'pcg', or 'fgmres'; "
                             f"got {gmg_outer_solver!r}.")
        if gmg_fine_smoother == "bf16" and not enable_fused_cuda:
            raise ValueError("SolverV4 BF16 fine smoothing requires enable_fused_cuda=True so the "
                             "WMMA path is available.")
        edof = np.asarray(_host(edof))
        self.ndof = int(ndof)
        self.E0 = float(E0)
        self.Emin = float(Emin)
        self.cg_tol = float(cg_tol)
        self.cg_maxiter = int(cg_maxiter)
        self._grid_dims = tuple(int(v) for v in (grid_dims or grid_from_edof(edof)))
        self._enable_warm_start = bool(enable_warm_start)
        self._enable_profiling = bool(enable_profiling)
        self._gmg_outer_solver = gmg_outer_solver
        self._gmg_restart = int(gmg_restart)
        self._free = np.asarray(_host(free), dtype=np.int64)
        self._n_free = len(self._free)
        self._n_elem = edof.shape[0]
        self._F_free_gpu = _as_dev(np.asarray(_host(F), dtype=np.float64)[self._free],
                                   torch.float64)
        self._free_gpu = torch.as_tensor(self._free.astype(np.int32), device="cuda")
        self._edof_gpu = torch.as_tensor(edof.astype(np.int32), device="cuda")
        self._KE_unit_gpu = torch.as_tensor(np.asarray(_host(KE_UNIT), dtype=np.float64),
                                            device="cuda")
        self._matfree_op = MatrixFreeKff(self._edof_gpu, self._KE_unit_gpu, self._free_gpu,
                                         self._n_free, self.ndof, grid=self._grid_dims)
        nx, ny, nz = self._grid_dims
        # the fused operator, for the bf16 fine smoother (solver_v4.py:111-119)
        self._fused_op = None
        if enable_fused_cuda:
            from .cuda_fused_matvec import FusedMatvec

            self._fused_op = FusedMatvec(self._edof_gpu, self._KE_unit_gpu, self.ndof,
                                         grid=self._grid_dims)
        self._matfree_gmg = GalerkinMatFreeGMG(
            mf_op=self._matfree_op, free=self._free, free_gpu=self._free_gpu, nelx=nx, nely=ny,
            nelz=nz, KE_UNIT=_host(KE_UNIT), n_levels=matfree_gmg_levels,
            coarse_smooth_iters=gmg_smooth_iters, fine_smoother=gmg_fine_smoother,
            fine_smoother_degree=gmg_fine_degree, smoother_type=gmg_smoother_type,
            cheb_lower_frac=gmg_cheb_lower_frac, level_precisions=gmg_level_precisions,
            cycle_type=gmg_cycle_type, fused_op=self._fused_op)
        self._u_prev = None
        self._prev_penal: Optional[float] = None
        self.last_outer_solver = ""
        self.last_cg_iters = 0
        self.last_cg_history: list = []
        self.last_true_rel_residual = float("nan")
        self.last_multigrid_quality: dict = {}
        self.last_timing: dict = {}

    def solve(self, rho_phys, penal: float):
        """(compliance, dc_phys) for the physical density (solver_v4.py:149-282)."""
        import torch

        t0 = time.perf_counter()
        if self._prev_penal is not None and abs(penal - self._prev_penal) > 0.5:
            self._u_prev = None
        self._prev_penal = penal
        rho = _as_dev(np.asarray(_host(rho_phys), dtype=np.float64).reshape(-1), torch.float64)
        E = self.Emin + (self.E0 - self.Emin) * rho ** penal
        fine_diag = self._matfree_op.extract_diagonal(E)
        self._matfree_gmg.setup(E, fine_diag=fine_diag)
        t_setup = time.perf_counter()
        A = self._matfree_op.operator(E)
        M = self._matfree_gmg.apply
        F = self._F_free_gpu
        x0 = self._u_prev if (self._enable_warm_start and self._u_prev is not None) else None
        outer = self._gmg_outer_solver
        if outer == "auto":
            outer = "fgmres" if self._matfree_gmg.uses_bf16_smoother() else "pcg"
        self.last_outer_solver = outer

        def run(start):
            hist: list = []
            if outer == "pcg":
                u, it, ok = _cupy_pcg(A, F, M, x0=start, tol=self.cg_tol,
                                      maxiter=self.cg_maxiter, history=hist)
            else:
                u, it, ok = _cupy_fgmres(A, F, M, x0=start, tol=self.cg_tol,
                                         maxiter=self.cg_maxiter, restart=self._gmg_restart,
                                         history=hist)
            return u, it, ok, hist

        u, it, ok, hist = run(x0)
        if not ok and x0 is not None:
            u, it, ok, hist = run(None)
        self.last_cg_iters = int(it)
        self.last_cg_history = list(hist)
        b_norm = nrm2(F)
        r = self._matfree_op.residual(F, u, E)
        self.last_true_rel_residual = nrm2(r) / max(b_norm, 1e-300)
        self.last_multigrid_quality = self._matfree_gmg.probe_quality(F)
        t_solve = time.perf_counter()
        if self._enable_warm_start:
            self._u_prev = u.clone()
        compliance = dot(F, u)
        dc = torch.empty_like(rho)
        _lib.call("gf_elem_sensitivity_f64", self._matfree_op.ctx_handle, _lib.ptr(u),
                  _lib.ptr(rho), float(penal), float(self.E0 - self.Emin), None, _lib.ptr(dc),
                  _lib.stream_handle())
        dc_host = dc.cpu().numpy()
        if self._enable_profiling:
            t_end = time.perf_counter()
            self.last_timing = {"gmg_setup_ms": (t_setup - t0) * 1e3,
                                "cg_ms": (t_solve - t_setup) * 1e3,
                                "sensitivity_ms": (t_end - t_solve) * 1e3,
                                "total_ms": (t_end - t0) * 1e3, "used_gmg": True,
                                "matrix_free": True, "outer_solver": outer, "cg_iters": int(it)}
        return compliance, dc_host


def _history_at(history, iteration: int) -> float:
    if not history:
        return float("inf")
    return float(history[min(max(iteration, 0), len(history) - 1)])


def _classify_probe(r50, r100, *, high_residual_threshold, plateau_residual_threshold,
                    plateau_ratio_threshold):
    if r50 >= high_residual_threshold:
        return True, "high_r50"
    ratio = r100 / max(r50, 1e-300)
    if r100 >= plateau_residual_threshold and ratio >= plateau_ratio_threshold:
        return True, "plateau"
    return False, "keep"


class GuardedAdaptiveFloorSolverV4(SolverV4):
    """SolverV4 with the guarded floor policy applied per SIMP iteration
    (run_simp_floor_trajectory.py:65-224)."""

    def __init__(self, *args, guarded_base_rho_min: float = 1e-12,
                 guarded_ladder=(1e-3, 1e-2), guarded_probe_iters: int = 100,
                 guarded_probe_r50: int = 50, guarded_high_residual_threshold: float = 1e-2,
                 guarded_plateau_residual_threshold: float = 1e-4,
                 guarded_plateau_ratio_threshold: float = 0.6, **kwargs):
        super().__init__(*args, **kwargs)
        self.guarded_base_rho_min = float(guarded_base_rho_min)
        self.guarded_ladder = tuple(float(v) for v in guarded_ladder)
        self.guarded_probe_iters = int(guarded_probe_iters)
        self.guarded_probe_r50 = int(guarded_probe_r50)
        self.guarded_high_residual_threshold = float(guarded_high_residual_threshold)
        self.guarded_plateau_residual_threshold = float(guarded_plateau_residual_threshold)
        self.guarded_plateau_ratio_threshold = float(guarded_plateau_ratio_threshold)
        self.policy_events: list = []
        self.last_selected_rho_min = self.guarded_base_rho_min
        self.last_policy_trigger = ""
        self.last_policy_fallback_used = 0
        self.last_policy_attempted_rho_mins = ""
        self.last_policy_probe_r50 = float("nan")
        self.last_policy_probe_r100 = float("nan")
        self.last_policy_probe_true_rel_residual = float("nan")

    def _solve_with_floor(self, rho_phys, penal, *, rho_min, maxiter, phase):
        old_emin, old_maxiter = self.Emin, self.cg_maxiter
        self.Emin = float(rho_min)
        self.cg_maxiter = int(maxiter)
        try:
            start = time.perf_counter()
            compliance, dc = SolverV4.solve(self, rho_phys, penal)
            metrics = {"phase": phase, "rho_min": float(rho_min), "maxiter": int(maxiter),
                       "compliance": float(compliance), "iters": int(self.last_cg_iters),
                       "true_rel_residual": float(self.last_true_rel_residual),
                       "history": list(self.last_cg_history),
                       "wall_time_s": time.perf_counter() - start}
            return compliance, dc, metrics
        finally:
            self.Emin, self.cg_maxiter = old_emin, old_maxiter

    def solve(self, rho_phys, penal: float):
        attempted = []
        compliance, dc, probe = self._solve_with_floor(
            rho_phys, penal, rho_min=self.guarded_base_rho_min,
            maxiter=self.guarded_probe_iters, phase="probe")
        attempted.append(probe)
        r50 = _history_at(probe["history"], self.guarded_probe_r50)
        r100 = _history_at(probe["history"], self.guarded_probe_iters)
        self.last_policy_probe_r50, self.last_policy_probe_r100 = r50, r100
        self.last_policy_probe_true_rel_residual = float(probe["true_rel_residual"])
        if probe["true_rel_residual"] <= self.cg_tol:
            selected, trigger, fallback_used = probe, "probe_converged_keep", 0
        else:
            raise_floor, trigger = _classify_probe(
                r50, r100, high_residual_threshold=self.guarded_high_residual_threshold,
                plateau_residual_threshold=self.guarded_plateau_residual_threshold,
                plateau_ratio_threshold=self.guarded_plateau_ratio_threshold)
            selected, fallback_used = None, 0
            if not raise_floor:
                compliance, dc, full = self._solve_with_floor(
                    rho_phys, penal, rho_min=self.guarded_base_rho_min,
                    maxiter=self.cg_maxiter, phase="guarded_keep_full")
                attempted.append(full)
                if full["true_rel_residual"] <= self.cg_tol:
                    selected = full
                else:
                    fallback_used = 1
                    trigger = f"{trigger}_guard_failed"
            if selected is None:
                for floor in self.guarded_ladder:
                    compliance, dc, lad = self._solve_with_floor(
                        rho_phys, penal, rho_min=floor, maxiter=self.cg_maxiter, phase="ladder")
                    attempted.append(lad)
                    if lad["true_rel_residual"] <= self.cg_tol:
                        selected = lad
                        break
            if selected is None:
                self.policy_events.append({
                    "trigger": trigger, "selected_rho_min": float("nan"), "fallback_used": 1,
                    "attempted_rho_mins": ";".join(f"{a['rho_min']:.0e}" for a in attempted),
                    "probe_r50": r50, "probe_r100": r100,
                    "accepted_true_rel_residual": float("nan")})
                raise RuntimeError("guarded adaptive floor policy exhausted the ladder")
        self.last_selected_rho_min = float(selected["rho_min"])
        self.last_true_rel_residual = float(selected["true_rel_residual"])
        self.last_cg_iters = int(selected["iters"])
        self.last_cg_history = list(selected["history"])
        self.last_policy_trigger = trigger
        self.last_policy_fallback_used = int(fallback_used)
        self.last_policy_attempted_rho_mins = ";".join(f"{a['rho_min']:.0e}" for a in attempted)
        self.policy_events.append({
            "trigger": trigger, "selected_rho_min": self.last_selected_rho_min,
            "fallback_used": self.last_policy_fallback_used,
            "attempted_rho_mins": self.last_policy_attempted_rho_mins,
            "probe_r50": r50, "probe_r100": r100,
            "probe_true_rel_residual": self.last_policy_probe_true_rel_residual,
            "accepted_true_rel_residual": self.last_true_rel_residual,
            "accepted_iters": self.last_cg_iters, "attempt_count": len(attempted),
            "policy_time_s": sum(float(a["wall_time_s"]) for a in attempted)})
        return compliance, dc
