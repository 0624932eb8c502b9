"""Oracle: density states and the guarded floor policy (NumPy).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Restates run_gmg_floor_detector_prospective.py:35-90 and 191-447.  Pinning
(``tests/test_oracle_policy.py``):

* ``bernoulli_mask`` + ``connectivity``: the reference tables' columns
  solid_count / n_components / largest_component for every (seed, p) of the
  72-state cantilever and 30-state bridge held-out suites (exact integers).
* ``classify``: recomputes the ``trigger`` column of every reference
  summary row from its own probe_r50 / probe_r100 columns (truth table).
* ``run_policy``: scripted solvers whose projected residual lies (the guard
  must reject and the ladder advance), short probe histories (r50/r100 index
  clamp), and the raised ladder stopping at the first guarded floor; the
  reference rows' own attempted-floor / fallback / false-keep columns are
  checked for the same ladder rules.
"""

from __future__ import annotations

import time
from collections import deque

import numpy as np


def bernoulli_field(seed: int, dims) -> np.ndarray:
    """default_rng(seed).random(n_e) reshaped (nelx, nely, nelz)
    (prospective:214-215)."""
    nx, ny, nz = dims
    return np.random.default_rng(seed).random(nx * ny * nz).reshape(nx, ny, nz)


def bernoulli_mask(seed: int, dims, p: float) -> np.ndarray:
    return bernoulli_field(seed, dims) < p


def connectivity(mask: np.ndarray) -> dict:
    """6-neighbour BFS statistics of the solid mask (prospective:35-74)."""
    nx, ny, nz = mask.shape
    comp = -np.ones(mask.shape, dtype=np.int32)
    nb = ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))
    ncomp = largest = s2l = 0
    for start in zip(*np.nonzero(mask)):
        if comp[start] >= 0:
            continue
        q = deque([start])
        comp[start] = ncomp
        size = 0
        sup = load = False
        while q:
            x, y, z = q.popleft()
            size += 1
            sup = sup or x == 0
            load = load or (x == nx - 1 and abs(y - ny // 2) <= 1 and abs(z - nz // 2) <= 1)
            for dx, dy, dz in nb:
                a, b, c = x + dx, y + dy, z + dz
                if 0 <= a < nx and 0 <= b < ny and 0 <= c < nz and mask[a, b, c] and comp[a, b, c] < 0:
                    comp[a, b, c] = ncomp
                    q.append((a, b, c))
        largest = max(largest, size)
        s2l += int(sup and load)
        ncomp += 1
    sc = int(mask.sum())
    return {"solid_count": sc, "solid_fraction": sc / mask.size, "n_components": ncomp,
            "largest_component": largest, "largest_fraction_of_solid": largest / max(sc, 1),
            "support_to_load_patch_components": s2l}


def classify(r50, r100, *, high_residual_threshold=1e-2, plateau_residual_threshold=1e-4,
             plateau_ratio_threshold=0.6):
    """(raise_floor, trigger) -- prospective:77-90."""
    if r50 >= high_residual_threshold:
        return True, "high_r50"
    ratio = r100 / max(r50, 1e-300)
    if r100 >= plateau_residual_threshold and ratio >= plateau_ratio_threshold:
        return True, "plateau"
    return False, "keep"


def simp_modulus(rho, rho_min, penal):
    """E = rho_min + (1 - rho_min) * rho**penal (prospective:221-222)."""
    return rho_min + (1.0 - rho_min) * np.asarray(rho, dtype=np.float64) ** penal


def run_policy(solid_flat, solve, F_free, A_op_factory, *, baseline_rho_min=1e-12,
               raised_rho_mins=(1e-3, 1e-2), penal=4.5, probe_iters=100, probe_r50=50,
               maxiter=300, tol=1e-6, high_residual_threshold=1e-2,
               plateau_residual_threshold=1e-4, plateau_ratio_threshold=0.6,
               severity_jump_r50_threshold=0.0, severity_jump_rho_min=1e-2,
               rho_field=None):
    """Probe -> classify -> ladder -> recomputed-residual guard -> fallback.

    ``solve(E, maxiter, history) -> (x, iters, reported)`` runs set-up + FGMRES;
    ``A_op_factory(E)`` returns the fine operator used by the guard.
    ``rho_field`` (continuous densities, clipped to [0, 1]) replaces the 0/1
    solid mask when given (run_gmg_floor_detector_density_field.py:170-181).
    Restates prospective:220-354 and returns the summary-row fields that do not
    depend on wall time, plus the histories.
    """
    b_norm = float(np.linalg.norm(F_free))
    raised = list(raised_rho_mins)

    def solve_at_floor(rho_min, mx, hist):
        if rho_field is None:
            rho = np.where(solid_flat, 1.0, rho_min)
        else:
            rho = np.clip(np.asarray(rho_field, dtype=np.float64), 0.0, 1.0)
        E = simp_modulus(rho, rho_min, penal)
        t0 = time.perf_counter()
        x, iters, reported = solve(E, mx, hist)
        final = float(np.linalg.norm(F_free - A_op_factory(E)(x)) / b_norm)
        return x, iters, bool(reported), final <= tol, final, time.perf_counter() - t0

    probe_hist = []
    _, _, _, _, _, t_probe = solve_at_floor(baseline_rho_min, probe_iters, probe_hist)
    r50 = probe_hist[min(probe_r50, len(probe_hist) - 1)]
    r100 = probe_hist[min(probe_iters, len(probe_hist) - 1)]
    raise_floor, trigger = classify(r50, r100, high_residual_threshold=high_residual_threshold,
                                    plateau_residual_threshold=plateau_residual_threshold,
                                    plateau_ratio_threshold=plateau_ratio_threshold)
    jump = raise_floor and severity_jump_r50_threshold > 0.0 and r50 >= severity_jump_r50_threshold
    if raise_floor and jump:
        ladder = [v for v in raised if v >= severity_jump_rho_min] or [severity_jump_rho_min]
    else:
        ladder = raised if raise_floor else [baseline_rho_min]
    st = {"attempted": [], "rho": ladder[-1], "iters": maxiter, "ok": False, "reported": False,
          "final": float("nan"), "x": None, "hist": [], "time": t_probe}

    def run_seq(seq):
        for rho_min in seq:
            if rho_min in st["attempted"]:
                continue
            st["attempted"].append(rho_min)
            h = []
            x, it, rep, ok, fin, t = solve_at_floor(rho_min, maxiter, h)
            st["hist"].append((rho_min, h))
            st.update(rho=rho_min, iters=it, ok=ok, reported=rep, final=fin, x=x)
            st["time"] += t
            if ok:
                break

    run_seq(ladder)
    fallback = False
    if (not st["ok"]) and (not raise_floor):
        fallback = True
        run_seq(raised)
    compliance = float(np.dot(F_free, st["x"])) if st["x"] is not None else float("nan")
    return {
        "probe_r50": r50, "probe_r100": r100, "probe_r100_over_r50": r100 / max(r50, 1e-300),
        "trigger": trigger, "predicted_raise_floor": int(raise_floor),
        "severity_jump_triggered": int(jump), "post_solve_fallback_triggered": int(fallback),
        "detector_false_keep": int((not raise_floor) and fallback),
        "recommended_rho_min": st["rho"], "attempted_rho_mins": list(st["attempted"]),
        "n_floor_attempts": len(st["attempted"]), "solve_converged": int(st["ok"]),
        "solver_reported_converged": int(st["reported"]), "solve_iters": st["iters"],
        "solve_final_rel_residual": st["final"], "compliance": compliance,
        "probe_history": probe_hist, "solve_history": st["hist"], "x": st["x"],
        "policy_time_s": st["time"],
    }


def filtered_sigmoid_density(seed: int, dims, radius=3.0, sharpness=8.0, volfrac=0.3):
    """Seeded N(0,1) field -> cone filter of the given radius (the reference's
    density-filter weights, pub_simp_solver.py:215-243) -> sigmoid(sharpness*(f-s))
    with the shift s bisected so that the mean equals ``volfrac``."""
    from scipy.ndimage import convolve

    nx, ny, nz = dims
    g = np.random.default_rng(seed).standard_normal(nx * ny * nz).reshape(nx, ny, nz)
    r = int(np.ceil(radius)) - 1
    ax = np.arange(-r, r + 1)
    dx, dy, dz = np.meshgrid(ax, ax, ax, indexing="ij")
    w = np.maximum(radius - np.sqrt(dx ** 2 + dy ** 2 + dz ** 2), 0.0)
    num = convolve(g, w, mode="constant", cval=0.0)
    den = convolve(np.ones_like(g), w, mode="constant", cval=0.0)
    f = num / den
    f = f / max(float(f.std()), 1e-30)
    lo, hi = float(f.min()) - 10.0, float(f.max()) + 10.0
    for _ in range(200):
        s = 0.5 * (lo + hi)
        m = float(np.mean(1.0 / (1.0 + np.exp(-sharpness * (f - s)))))
        if m > volfrac:
            lo = s
        else:
            hi = s
    s = 0.5 * (lo + hi)
    return (1.0 / (1.0 + np.exp(-sharpness * (f - s)))).ravel()
