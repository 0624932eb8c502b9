"""Guarded floor-selection policy on the HIP solver stack.

Library form of the reference's script-level policy
(run_gmg_floor_detector_prospective.py:35-515, density-field variant
run_gmg_floor_detector_density_field.py:170-299, component builder
run_experiments_e1_e10.py:128-179):

  probe FGMRES (probe_iters) at the baseline floor
  -> classify(r50, r100)  (high_r50 | plateau | keep)
  -> floor ladder, each attempt accepted only if the RECOMPUTED relative
     residual ||F - K(E) x|| / ||F|| <= tol
  -> post-solve fallback to the raised ladder when a kept floor fails.

Each state produces a row with the reference's summary column names
(prospective:356-416); device-memory columns report the HIP allocator.
"""

from __future__ import annotations

import time
from collections import deque

import numpy as np

from . import _lib
from .bc_generator import generate_bc
from .multigrid_v4 import GalerkinMatFreeGMG
from .presets import cantilever, get_preset
from .pub_simp_solver import KE_UNIT_3D, _edof_table_3d
from .solver_v2 import MatrixFreeKff, nrm2

CANONICAL_STACK = dict(
    n_levels=4, fine_smoother="fp64", smoother_type="chebyshev", cycle_type="v",
    coarse_correction_policy="residual_line_search",
    root_local_correction_mode="node_block_inner_fgmres",
    root_local_correction_inner_steps=20, root_local_correction_inner_restart=20,
    inner_krylov_steps=10, inner_krylov_restart=10,
)


def _classify(r50: float, r100: float, *, high_residual_threshold: float,
              plateau_residual_threshold: float, plateau_ratio_threshold: float):
    """(raise_floor, trigger) -- prospective:77-90."""
    if r50 >= high_residual_threshold:
        return True, "high_r50"
    ratio = r100 / max(r50, 1e-300)
    if r100 >= plateau_residual_threshold and ratio >= plateau_ratio_threshold:
        return True, "plateau"
    return False, "keep"


classify = _classify


def _connectivity(mask: np.ndarray) -> dict:
    """6-neighbour component statistics of the solid mask (prospective:35-74)."""
    nx, ny, nz = mask.shape
    lab = np.full(mask.shape, -1, dtype=np.int32)
    ncomp = largest = s2l = 0
    steps = ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))
    for start in zip(*np.nonzero(mask)):
        if lab[start] >= 0:
            continue
        lab[start] = ncomp
        q = deque([start])
        size = 0
        sup = load = False
        while q:
            x, y, z = q.popleft()
            size += 1
            sup |= x == 0
            load |= (x == nx - 1 and abs(y - ny // 2) <= 1 and abs(z - nz // 2) <= 1)
            for dx, dy, dz in steps:
                a, b, c = x + dx, y + dy, z + dz
                if 0 <= a < nx and 0 <= b < ny and 0 <= c < nz and mask[a, b, c] and lab[a, b, c] < 0:
                    lab[a, b, c] = ncomp
                    q.append((a, b, c))
        largest = max(largest, size)
        s2l += int(sup and load)
        ncomp += 1
    sc = int(mask.sum())
    return {"solid_count": sc, "solid_fraction": sc / mask.size, "n_components": ncomp,
            "largest_component": largest, "largest_fraction_of_solid": largest / max(sc, 1),
            "support_to_load_patch_components": s2l}


connectivity = _connectivity


def _build_components(preset_or_spec, *, fused_op=None, **stack):
    """(spec, bc, free, free_gpu, edof_gpu, F_free_gpu, mf_op, gmg) on the HIP
    stack (run_experiments_e1_e10.py:128-179)."""
    import torch

    _lib.require_gpu()
    spec = get_preset(preset_or_spec) if isinstance(preset_or_spec, str) else preset_or_spec
    bc = generate_bc(spec)
    free = bc.free_dofs.astype(np.int32)
    edof = _edof_table_3d(spec.nelx, spec.nely, spec.nelz)
    edof_gpu = torch.as_tensor(edof, device="cuda")
    free_gpu = torch.as_tensor(free, device="cuda")
    F_free_gpu = torch.as_tensor(bc.F[free].astype(np.float64), device="cuda")
    KE = torch.as_tensor(KE_UNIT_3D, device="cuda")
    mf_op = MatrixFreeKff(edof_gpu, KE, free_gpu, len(free), bc.ndof,
                          grid=(spec.nelx, spec.nely, spec.nelz))
    if fused_op is None and stack.get("fine_smoother") == "bf16":
        # the bf16 smoother runs the fused operator's bf16 form, which the
        # reference builds with enable_fused_cuda (solver_v4.py:111-119)
        from .cuda_fused_matvec import FusedMatvec

        fused_op = FusedMatvec(edof_gpu, KE, bc.ndof, grid=(spec.nelx, spec.nely, spec.nelz))
    gmg = GalerkinMatFreeGMG(mf_op=mf_op, free=free, free_gpu=free_gpu, nelx=spec.nelx,
                             nely=spec.nely, nelz=spec.nelz, KE_UNIT=KE_UNIT_3D,
                             fused_op=fused_op, **stack)
    return spec, bc, free, free_gpu, edof_gpu, F_free_gpu, mf_op, gmg


def _sync():
    """Wait for this thread's current stream only (other threads may be
    driving independent solves on their own streams)."""
    import torch

    torch.cuda.current_stream().synchronize()


def _mem_used_mb() -> float:
    import torch

    free_b, total_b = torch.cuda.mem_get_info()
    return (total_b - free_b) / 1e6


class GuardedFloorPolicy:
    """One component set (operator + hierarchy) reused across states, so the
    hierarchy's cached spectral estimate follows the reference's session
    semantics (multigrid_v4.py:1170-1192)."""

    def __init__(self, preset_or_spec, *, penal=4.5, baseline_rho_min=1e-12,
                 raised_rho_mins=(1e-3, 1e-2), probe_iters=100, probe_r50=50, restart=300,
                 maxiter=300, tol=1e-6, high_residual_threshold=1e-2,
                 plateau_residual_threshold=1e-4, plateau_ratio_threshold=0.6,
                 severity_jump_r50_threshold=0.0, severity_jump_rho_min=1e-2,
                 stack_variant="canonical", preset_name=None, reuse_probe=True, **stack):
        st = dict(CANONICAL_STACK)
        st.update(stack)
        self.stack = st
        self.stack_variant = stack_variant
        (self.spec, self.bc, self.free, self.free_gpu, self.edof_gpu, self.F, self.mf_op,
         self.gmg) = _build_components(preset_or_spec, **st)
        self.preset = preset_name or (preset_or_spec if isinstance(preset_or_spec, str) else "custom")
        self.penal = float(penal)
        self.baseline_rho_min = float(baseline_rho_min)
        self.raised_rho_mins = [float(v) for v in raised_rho_mins]
        self.probe_iters = int(probe_iters)
        self.probe_r50 = int(probe_r50)
        self.restart = int(restart)
        self.maxiter = int(maxiter)
        self.tol = float(tol)
        self.thr = dict(high_residual_threshold=float(high_residual_threshold),
                        plateau_residual_threshold=float(plateau_residual_threshold),
                        plateau_ratio_threshold=float(plateau_ratio_threshold))
        self.severity_jump_r50_threshold = float(severity_jump_r50_threshold)
        self.severity_jump_rho_min = float(severity_jump_rho_min)
        # a kept baseline floor continues the probe's Arnoldi cycle instead of
        # repeating its set-up and first probe_iters steps (identical result)
        self.reuse_probe = bool(reuse_probe)
        self.b_norm = nrm2(self.F)
        self.n_elem = self.spec.nelx * self.spec.nely * self.spec.nelz

    @classmethod
    def from_cantilever(cls, dims, *, load="tip", **kw):
        spec = cantilever(*dims, load=load)
        return cls(spec, preset_name=f"cantilever_{dims[0]}x{dims[1]}x{dims[2]}_{load}", **kw)

    # -------------------------------------------------------------------
    # True for a policy whose first set-up of a state may be at another
    # floor than the probe's (the learned policy's proposals): a session then
    # primes its first instance with the probe's moduli before that state
    primes_first_state = False

    def first_moduli(self, state: dict) -> np.ndarray:
        """Moduli of the plain policy's first set-up on ``state`` (the probe at
        the baseline floor).  A session sets up every instance with them
        first, so its cached spectral estimate is the plain session's
        whichever policy it runs."""
        return self.moduli(self.baseline_rho_min, state.get("solid"), state.get("rho_field"))

    def moduli(self, rho_min: float, solid=None, rho_field=None) -> np.ndarray:
        """E = rho_min + (1 - rho_min) rho^p on the host: rho = 1 on the solid
        mask and rho_min elsewhere (prospective:216-222), or a continuous
        density clipped to [0, 1] (density_field:170-181)."""
        if rho_field is not None:
            rho = np.clip(np.asarray(rho_field, dtype=np.float64).reshape(-1), 0.0, 1.0)
        else:
            rho = np.where(np.asarray(solid).reshape(-1), 1.0, rho_min)
        return rho_min + (1.0 - rho_min) * rho ** self.penal

    def solve_at_floor(self, rho_min, *, maxiter, history, solid=None, rho_field=None):
        import torch

        E_host = self.moduli(rho_min, solid, rho_field)
        E = torch.as_tensor(E_host, device="cuda")
        _sync()
        mem_before = _mem_used_mb()
        t0 = time.perf_counter()
        self.gmg.setup(E)
        _sync()
        setup_time = time.perf_counter() - t0
        mem_setup = _mem_used_mb()
        t1 = time.perf_counter()
        x, iters, reported = self.gmg.solve(self.F, tol=self.tol, maxiter=maxiter,
                                            restart=self.restart, history=history)
        _sync()
        solve_time = time.perf_counter() - t1
        r = self.mf_op.residual(self.F, x, E)
        final = nrm2(r) / self.b_norm
        metrics = {"setup_time_s": setup_time, "solve_time_s": solve_time,
                   "gpu_used_mb_before": mem_before, "gpu_used_mb_after_setup": mem_setup,
                   "gpu_used_mb_after_solve": _mem_used_mb()}
        return E, x, iters, bool(reported), final <= self.tol, final, metrics

    def continue_probe(self, E, probe_hist, *, maxiter, history):
        """The kept baseline floor's solve as the continuation of the probe.

        The reference solves the kept floor afresh (set-up at the same moduli,
        FGMRES from x = 0 with restart >= probe_iters: prospective:263-277 then
        313-351), so its first probe_iters Arnoldi steps repeat the probe's.
        Every kernel here is deterministic and the set-up at the same moduli
        reproduces the same hierarchy, so continuing the probe's cycle to
        ``maxiter`` gives the from-scratch solve bit for bit (history, iterate,
        iteration count) without repeating it.  Raises ContinueUnavailable
        when the probe cannot be continued."""
        t1 = time.perf_counter()
        new: list = []
        x, iters, reported = self.gmg.solve_continue(self.F, tol=self.tol, maxiter=maxiter,
                                                     restart=self.restart, history=new)
        _sync()
        solve_time = time.perf_counter() - t1
        history.extend(probe_hist)
        history.extend(new)
        r = self.mf_op.residual(self.F, x, E)
        final = nrm2(r) / self.b_norm
        metrics = {"setup_time_s": 0.0, "solve_time_s": solve_time,
                   "gpu_used_mb_before": float("nan"), "gpu_used_mb_after_setup": float("nan"),
                   "gpu_used_mb_after_solve": _mem_used_mb(), "continued_probe": True}
        return E, x, iters, bool(reported), final <= self.tol, final, metrics

    def run_state(self, solid=None, *, rho_field=None, seed=None, probability=None,
                  with_connectivity=False, keep_x=True):
        """Probe -> classify -> ladder -> guard -> fallback for one state."""
        from .multigrid_v4 import ContinueUnavailable
        from .solver_v2 import dot

        probe_hist: list = []
        probe_E, _, _, _, _, _, probe_m = self.solve_at_floor(
            self.baseline_rho_min, maxiter=self.probe_iters, history=probe_hist, solid=solid,
            rho_field=rho_field)
        r50 = probe_hist[min(self.probe_r50, len(probe_hist) - 1)]
        r100 = probe_hist[min(self.probe_iters, len(probe_hist) - 1)]
        raise_floor, trigger = _classify(r50, r100, **self.thr)
        jump = (raise_floor and self.severity_jump_r50_threshold > 0.0
                and r50 >= self.severity_jump_r50_threshold)
        if raise_floor and jump:
            ladder = [v for v in self.raised_rho_mins if v >= self.severity_jump_rho_min]
            ladder = ladder or [self.severity_jump_rho_min]
        else:
            ladder = self.raised_rho_mins if raise_floor else [self.baseline_rho_min]
        st = {"attempted": [], "rho": ladder[-1], "iters": self.maxiter, "ok": False,
              "reported": False, "final": float("nan"), "x": None, "E": None, "hist": [],
              "sel": None, "fail_setup": 0.0, "fail_solve": 0.0}

        def run_seq(seq):
            for rho_min in seq:
                if rho_min in st["attempted"]:
                    continue
                st["attempted"].append(rho_min)
                h: list = []
                res = None
                if (self.reuse_probe and len(st["attempted"]) == 1
                        and rho_min == self.baseline_rho_min):
                    try:
                        res = self.continue_probe(probe_E, probe_hist, maxiter=self.maxiter,
                                                  history=h)
                    except ContinueUnavailable:
                        res = None
                if res is None:
                    res = self.solve_at_floor(rho_min, maxiter=self.maxiter, history=h,
                                              solid=solid, rho_field=rho_field)
                E, x, it, rep, ok, fin, m = res
                st["hist"].append((rho_min, h))
                st.update(rho=rho_min, iters=it, ok=ok, reported=rep, final=fin, x=x, E=E, sel=m)
                if ok:
                    break
                st["fail_setup"] += m["setup_time_s"]
                st["fail_solve"] += m["solve_time_s"]

        run_seq(ladder)
        fallback = False
        if (not st["ok"]) and (not raise_floor):
            fallback = True
            run_seq(self.raised_rho_mins)
        sel = st["sel"] or {k: float("nan") for k in ("setup_time_s", "solve_time_s",
                                                      "gpu_used_mb_after_solve")}
        compliance = dot(self.F, st["x"]) if st["x"] is not None else float("nan")
        row = {
            "preset": self.preset, "seed": seed, "solid_probability": probability,
            "baseline_rho_min": self.baseline_rho_min,
            "raised_rho_min": self.raised_rho_mins[0],
            "raised_rho_mins": ";".join(f"{v:.0e}" for v in self.raised_rho_mins),
            "recommended_rho_min": st["rho"],
            "attempted_rho_mins": ";".join(f"{v:.0e}" for v in st["attempted"]),
            "n_floor_attempts": len(st["attempted"]), "penal": self.penal,
            "probe_iters": self.probe_iters, "restart": self.restart, "maxiter": self.maxiter,
            "tol": self.tol, "stack_variant": self.stack_variant,
            **{k: self.stack[k] for k in ("n_levels", "fine_smoother", "smoother_type",
                                          "cycle_type", "coarse_correction_policy",
                                          "root_local_correction_mode",
                                          "root_local_correction_inner_steps",
                                          "root_local_correction_inner_restart",
                                          "inner_krylov_steps", "inner_krylov_restart")},
            "probe_r50": r50, "probe_r100": r100, "probe_r100_over_r50": r100 / max(r50, 1e-300),
            "probe_setup_time_s": probe_m["setup_time_s"],
            "probe_solve_time_s": probe_m["solve_time_s"],
            "probe_gpu_used_mb_after_solve": probe_m["gpu_used_mb_after_solve"],
            "trigger": trigger, "predicted_raise_floor": int(raise_floor),
            "severity_jump_triggered": int(jump),
            "severity_jump_r50_threshold": self.severity_jump_r50_threshold,
            "severity_jump_rho_min": self.severity_jump_rho_min,
            "post_solve_fallback_triggered": int(fallback),
            "detector_false_keep": int((not raise_floor) and fallback),
            "solve_converged": int(st["ok"]), "solver_reported_converged": int(st["reported"]),
            "solve_iters": st["iters"], "solve_final_rel_residual": st["final"],
            "selected_setup_time_s": sel["setup_time_s"],
            "selected_solve_time_s": sel["solve_time_s"],
            "selected_gpu_used_mb_after_solve": sel["gpu_used_mb_after_solve"],
            "failed_ladder_setup_time_s": st["fail_setup"],
            "failed_ladder_solve_time_s": st["fail_solve"],
            "recorded_policy_time_s": (probe_m["setup_time_s"] + probe_m["solve_time_s"]
                                       + st["fail_setup"] + st["fail_solve"]
                                       + sel["setup_time_s"] + sel["solve_time_s"]),
            "compliance": compliance,
            "selected_continued_probe": int(bool(sel.get("continued_probe", False))),
        }
        if with_connectivity and solid is not None:
            m = np.asarray(solid).reshape(self.spec.nelx, self.spec.nely, self.spec.nelz)
            row.update(_connectivity(m))
        row["probe_history"] = probe_hist
        row["solve_history"] = st["hist"]
        if keep_x:
            row["x"] = st["x"]
        return row


class PolicySession:
    """Many states through the guarded policy as ONE reference session, with
    several solver instances running concurrently.

    The reference runs a multi-state evaluation as one session: one hierarchy
    whose lambda_max is estimated at the session's first set-up (the first
    state's probe) and reused while max(E) stays within 10 %
    (multigrid_v4.py:1170-1192).  Here ``workers`` independent
    ``GuardedFloorPolicy`` instances each own a HIP stream and are driven from
    host threads (the C calls release the GIL).  The first state runs alone on
    instance 0; every other instance then performs the same first set-up on
    that state's probe moduli, so all instances hold the identical cached
    lambda.  Every kernel is deterministic, so each state's row equals what a
    single sequential session would produce, whatever the interleaving.
    """

    def __init__(self, preset_or_spec, *, workers: int = 4, runner_cls=None, **kw):
        import torch

        if workers < 1:
            raise ValueError("workers must be >= 1")
        cls = runner_cls or GuardedFloorPolicy
        self.runners = [cls(preset_or_spec, **kw) for _ in range(workers)]
        self.streams = [torch.cuda.Stream() for _ in range(workers)]
        self._started = False

    @property
    def workers(self) -> int:
        return len(self.runners)

    def _run_one(self, k, st):
        import torch

        with torch.cuda.stream(self.streams[k]):
            row = self.runners[k].run_state(st.get("solid"), rho_field=st.get("rho_field"),
                                            seed=st.get("seed"),
                                            probability=st.get("probability"),
                                            keep_x=st.get("keep_x", False),
                                            **st.get("extra", {}))
            _sync()
        return row

    def run(self, states, on_done=None):
        """states: list of dicts with ``solid`` (or ``rho_field``) and optional
        ``seed`` / ``probability`` / ``keep_x`` / ``extra`` (keyword arguments
        passed on to the runner's ``run_state``).  Returns the rows in order.
        ``on_done(i)`` (optional) is called as each state finishes, from the
        thread that ran it."""
        import queue
        import threading

        import torch

        rows = [None] * len(states)
        start = 0
        if not self._started and states:
            st0 = states[0]
            E0 = torch.as_tensor(self.runners[0].first_moduli(st0), device="cuda")
            if self.runners[0].primes_first_state:
                with torch.cuda.stream(self.streams[0]):
                    self.runners[0].gmg.setup(E0)
                    _sync()
            rows[0] = self._run_one(0, st0)
            if on_done is not None:
                on_done(0)
            for k in range(1, self.workers):
                with torch.cuda.stream(self.streams[k]):
                    self.runners[k].gmg.setup(E0)
                    _sync()
            self._started = True
            start = 1
        todo = queue.Queue()
        for i in range(start, len(states)):
            todo.put(i)
        errors = []

        def worker(k):
            while not errors:
                try:
                    i = todo.get_nowait()
                except queue.Empty:
                    return
                try:
                    rows[i] = self._run_one(k, states[i])
                    if on_done is not None:
                        on_done(i)
                except BaseException as ex:  # surfaced in the caller
                    errors.append(ex)
                    return

        threads = [threading.Thread(target=worker, args=(k,), daemon=True)
                   for k in range(min(self.workers, max(1, len(states) - start)))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        return rows


SUMMARY_FIELDS = [
    "preset", "seed", "solid_probability", "baseline_rho_min", "raised_rho_min",
    "raised_rho_mins", "recommended_rho_min", "attempted_rho_mins", "n_floor_attempts", "penal",
    "probe_iters", "restart", "maxiter", "tol", "stack_variant", "n_levels", "fine_smoother",
    "smoother_type", "cycle_type", "coarse_correction_policy", "root_local_correction_mode",
    "root_local_correction_inner_steps", "root_local_correction_inner_restart",
    "inner_krylov_steps", "inner_krylov_restart", "solid_count", "solid_fraction",
    "n_components", "largest_component", "largest_fraction_of_solid",
    "support_to_load_patch_components", "probe_r50", "probe_r100", "probe_r100_over_r50",
    "probe_setup_time_s", "probe_solve_time_s", "probe_gpu_used_mb_after_solve", "trigger",
    "predicted_raise_floor", "severity_jump_triggered", "severity_jump_r50_threshold",
    "severity_jump_rho_min", "post_solve_fallback_triggered", "detector_false_keep",
    "solve_converged", "solver_reported_converged", "solve_iters", "solve_final_rel_residual",
    "selected_setup_time_s", "selected_solve_time_s", "selected_gpu_used_mb_after_solve",
    "failed_ladder_setup_time_s", "failed_ladder_solve_time_s", "recorded_policy_time_s",
    "compliance",
]
