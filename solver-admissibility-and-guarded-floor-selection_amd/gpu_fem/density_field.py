"""Guarded floor policy over continuous density fields, with the reference's
CSV schema.

Counterpart of experiments/phase5/run_gmg_floor_detector_density_field.py
(the density-field driver, :106-420): each density (one value per element,
clipped to [0, 1]) goes through probe -> classify -> floor ladder ->
recomputed-residual guard -> post-solve fallback with E = rho_min +
(1 - rho_min) * density**penal, in ONE session (one hierarchy, so the
spectral estimate is cached from the first set-up as in the reference).
Connectivity statistics use the solid mask density >= solid_threshold.

Also provides the synthetic filtered-sigmoid density used for the 1M-element
precision configuration (seeded N(0,1) noise -> cone filter -> sigmoid ->
shift to a target volume fraction) and CSV writers for summary and history
rows of both the density-field and the Bernoulli drivers.
"""

from __future__ import annotations

import csv
import os
from typing import Iterable, Optional, Sequence

import numpy as np

from .policy import GuardedFloorPolicy, _connectivity

# run_gmg_floor_detector_density_field.py:386-432 (column order of the
# reference's density_detector_summary.csv, kept for drop-in output)
DENSITY_SUMMARY_FIELDS = [
    "preset", "density_path", "density_name", "n_elem", "density_min", "density_max",
    "density_mean", "density_grayness", "baseline_rho_min", "raised_rho_mins",
    "recommended_rho_min", "attempted_rho_mins", "n_floor_attempts", "penal", "fine_smoother",
    "probe_iters", "restart", "maxiter", "tol", "solid_count", "solid_fraction", "n_components",
    "largest_component", "largest_fraction_of_solid", "probe_r50", "probe_r100",
    "probe_r100_over_r50", "probe_setup_time_s", "probe_solve_time_s",
    "probe_gpu_used_mb_after_solve", "probe_cupy_pool_used_mb_after_solve", "trigger",
    "predicted_raise_floor", "post_solve_fallback_triggered", "detector_false_keep",
    "solve_converged", "solver_reported_converged", "solve_iters", "solve_final_rel_residual",
    "selected_setup_time_s", "selected_solve_time_s", "selected_gpu_used_mb_after_solve",
    "selected_cupy_pool_used_mb_after_solve", "selected_cupy_pool_total_mb_after_solve",
    "failed_ladder_setup_time_s", "failed_ladder_solve_time_s", "recorded_policy_time_s",
]
DENSITY_HISTORY_FIELDS = ["density_name", "density_path", "phase", "rho_min", "iter",
                          "rel_residual"]
# run_gmg_floor_detector_prospective.py history schema
BERNOULLI_HISTORY_FIELDS = ["seed", "solid_probability", "phase", "rho_min", "iter",
                            "rel_residual"]


def write_csv(path: str, rows: Sequence[dict], fieldnames: Sequence[str]) -> None:
    """CSV with the given column order; keys outside it are ignored and
    missing keys are left empty (as csv.DictWriter does)."""
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(fieldnames), extrasaction="ignore")
        w.writeheader()
        w.writerows(rows)


def history_rows(row: dict, key: dict, baseline_rho_min: float) -> list:
    """Per-iteration rows of one policy row: the probe at the baseline floor,
    then every attempted solve (density_field.py:354-376)."""
    out = []
    for i, v in enumerate(row["probe_history"]):
        out.append(dict(key, phase="probe", rho_min=baseline_rho_min, iter=i,
                        rel_residual=float(v)))
    for rho_min, hist in row["solve_history"]:
        for i, v in enumerate(hist):
            out.append(dict(key, phase="recommended_solve", rho_min=rho_min, iter=i,
                            rel_residual=float(v)))
    return out


def _torch_pool_mb():
    import torch

    return (torch.cuda.memory_allocated() / 1e6, torch.cuda.memory_reserved() / 1e6)


def run_density_fields(densities: Iterable, *, preset="cantilever_gpu_medium",
                       names: Optional[Sequence[str]] = None,
                       paths: Optional[Sequence[str]] = None, penal: float = 4.5,
                       solid_threshold: float = 0.5, fine_smoother: str = "fp64",
                       baseline_rho_min: float = 1e-12, raised_rho_mins=(1e-3, 1e-2),
                       probe_iters: int = 100, probe_r50: int = 50, restart: int = 300,
                       maxiter: int = 300, tol: float = 1e-6,
                       high_residual_threshold: float = 1e-2,
                       plateau_residual_threshold: float = 1e-4,
                       plateau_ratio_threshold: float = 0.6, runner=None, **stack):
    """Run the guarded policy on each density field (arrays or .npy paths) in
    one session.  Returns (summary_rows, history_rows) in the reference's
    schema.  ``runner`` may pass an existing GuardedFloorPolicy (its session
    state, i.e. the cached spectral estimate, is then continued)."""
    if runner is None:
        runner = GuardedFloorPolicy(
            preset, penal=penal, baseline_rho_min=baseline_rho_min,
            raised_rho_mins=raised_rho_mins, probe_iters=probe_iters, probe_r50=probe_r50,
            restart=restart, maxiter=maxiter, tol=tol,
            high_residual_threshold=high_residual_threshold,
            plateau_residual_threshold=plateau_residual_threshold,
            plateau_ratio_threshold=plateau_ratio_threshold, fine_smoother=fine_smoother,
            **stack)
    spec = runner.spec
    n_elem = spec.nelx * spec.nely * spec.nelz
    summary, hist = [], []
    for k, dens in enumerate(densities):
        path = ""
        if isinstance(dens, (str, os.PathLike)):
            path = str(dens)
            dens = np.load(path)
        elif paths is not None:
            path = str(paths[k])
        d = np.asarray(dens, dtype=np.float64).reshape(-1)
        if d.size != n_elem:
            raise ValueError(f"density {k} has {d.size} entries, expected {n_elem} for the grid")
        d = np.clip(d, 0.0, 1.0)
        name = (names[k] if names is not None
                else (os.path.basename(os.path.dirname(path)) if path else f"density_{k}"))
        mask = d.reshape(spec.nelx, spec.nely, spec.nelz) >= solid_threshold
        row = runner.run_state(rho_field=d, keep_x=False)
        alloc_mb, total_mb = _torch_pool_mb()
        row.update(_connectivity(mask))
        row.update({
            "preset": runner.preset, "density_path": path, "density_name": name,
            "n_elem": n_elem, "density_min": float(d.min()), "density_max": float(d.max()),
            "density_mean": float(d.mean()),
            "density_grayness": float(np.mean(4.0 * d * (1.0 - d))),
            "fine_smoother": runner.stack.get("fine_smoother", fine_smoother),
            "probe_cupy_pool_used_mb_after_solve": alloc_mb,
            "selected_cupy_pool_used_mb_after_solve": alloc_mb,
            "selected_cupy_pool_total_mb_after_solve": total_mb,
        })
        summary.append(row)
        hist.extend(history_rows(row, {"density_name": name, "density_path": path},
                                 runner.baseline_rho_min))
    return summary, hist


def filtered_sigmoid_density(seed: int, dims, radius: float = 3.0, sharpness: float = 8.0,
                             volfrac: float = 0.3) -> np.ndarray:
    """Synthetic 'optimised-like' density (the 1M-element precision
    configuration): default_rng(seed).standard_normal per element, a cone
    filter of the given radius with the reference's density-filter weights
    w = max(0, r - dist) (pub_simp_solver.py:215-243) normalised per element,
    standardised, then sigmoid(sharpness * (f - s)) with s found by bisection
    so that the mean density equals volfrac.  Flat array, element order."""
    from scipy.ndimage import correlate

    nx, ny, nz = (int(v) for v in dims)
    g = np.random.default_rng(seed).standard_normal(nx * ny * nz).reshape(nx, ny, nz)
    half = int(np.ceil(radius)) - 1
    off = np.arange(-half, half + 1)
    ox, oy, oz = np.meshgrid(off, off, off, indexing="ij")
    w = np.maximum(radius - np.sqrt(ox ** 2 + oy ** 2 + oz ** 2), 0.0)
    num = correlate(g, w, mode="constant", cval=0.0)
    den = correlate(np.ones_like(g), w, mode="constant", cval=0.0)
    f = num / den
    f = f / max(float(f.std()), 1e-30)
    lo, hi = float(f.min()) - 10.0, float(f.max()) + 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        mean = float(np.mean(1.0 / (1.0 + np.exp(-sharpness * (f - mid)))))
        if mean > volfrac:
            lo = mid
        else:
            hi = mid
    s = 0.5 * (lo + hi)
    return (1.0 / (1.0 + np.exp(-sharpness * (f - s)))).ravel()
