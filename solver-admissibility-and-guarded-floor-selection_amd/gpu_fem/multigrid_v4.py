"""Galerkin matrix-free GMG preconditioner and flexible GMRES over HIP kernels.

Drop-in counterparts of multigrid_v4.py of the reference:
  GalerkinMatFreeGMG  multigrid_v4.py:537-2024 (production path)
  _cupy_fgmres        multigrid_v4.py:416-534
The hierarchy, its set-up and the whole nonlinear preconditioner (inner
FGMRES over V/W-cycles, safeguarded coarse correction, root-local node-block
FGMRES) run inside libgfhip.so without host synchronisation.  When the outer
FGMRES is called with ``A_op = mf_op.operator(E)`` (or the same moduli the
hierarchy was set up with) and ``M_op = gmg.apply`` it runs natively as well
(``gf_gmg_solve``); any other callables take the generic path below.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import asdict, dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from .solver_v2 import FineOperator, MatrixFreeKff, _as_dev, _host, axpby, dot, nrm2

_PREC = {"fp64": 0, "fp32": 1, "bf16": 2}
_PREC_NAME = {0: "fp64", 1: "fp32", 2: "bf16"}
_ENRICH = {"none": 0, "weighted_residual_second": 1, "weighted_gap_second": 2,
           "void_residual_second": 3, "void_gap_second": 4}
_ROOT_LOCAL = {"none": 0, "node_block_line_search": 1, "node_block_inner_fgmres": 2}


@dataclass
class MultigridLevelStats:
    level: int
    kind: str
    n_elem: int
    n_free: int
    nnz: int
    estimated_vram_bytes: int


@dataclass
class MultigridCycleVisit:
    """One level visit of a traced cycle (multigrid_v4.py:39-77; same fields)."""
    visit_id: int
    level: int
    cycle_type: str
    is_coarsest: bool
    n_coarse_corrections: int
    rhs_norm: float
    jacobi_action_norm: float
    presmooth_solution_norm: float
    presmooth_residual_norm: float
    presmooth_residual_ratio: float
    restricted_rhs_norm: float
    restricted_rhs_ratio: float
    projector_norm: float
    coarse_space_gap_norm: float
    coarse_space_gap_ratio: float
    coarse_correction_norm: float
    prolongated_correction_norm: float
    correction_alignment: float
    corrected_residual_norm: float
    corrected_residual_vs_presmooth: float
    correction_scale_raw: Optional[float] = None
    correction_scale_applied: float = 1.0
    second_restricted_rhs_norm: Optional[float] = None
    second_coarse_correction_norm: Optional[float] = None
    second_prolongated_correction_norm: Optional[float] = None
    second_correction_alignment: Optional[float] = None
    second_correction_scale_raw: Optional[float] = None
    second_correction_scale_applied: Optional[float] = None
    second_corrected_residual_norm: Optional[float] = None
    second_corrected_residual_vs_corrected: Optional[float] = None
    postsmooth_solution_norm: float = 0.0
    postsmooth_residual_norm: float = 0.0
    postsmooth_residual_ratio: float = 0.0
    coarsest_mode: str = ""
    notes: str = ""

    def to_dict(self) -> dict:
        return asdict(self)


@dataclass
class MultigridCycleTrace:
    """A traced cycle plus the root-local correction (multigrid_v4.py:80-112)."""
    cycle_type: str
    rhs_norm: float
    jacobi_action_norm: float
    output_norm: float
    z_over_jacobi: float
    rz: float
    pd: bool
    output_residual_norm: float
    output_residual_ratio: float
    contrast_ratio: float
    is_high_contrast: bool
    fine_degree_adapt: int
    coarse_iters_adapt: int
    coarse_correction_policy: str
    coarse_correction_acceptance_ratio: float
    coarse_correction_max_scale: float
    root_enrichment_mode: str
    root_enrichment_weight_floor: float
    root_enrichment_weight_power: float
    root_local_correction_mode: str
    root_local_correction_inner_steps: int
    root_local_correction_inner_restart: int
    root_local_correction_raw_scale: Optional[float]
    root_local_correction_applied_scale: float
    root_local_correction_residual_ratio: float
    inner_krylov_steps: int
    inner_krylov_restart: int
    visits: list = field(default_factory=list)

    def to_dict(self) -> dict:
        return asdict(self)


def _safe_ratio(num: float, den: float) -> float:
    """multigrid_v4.py:840-843."""
    if abs(den) <= 1e-300:
        return 0.0 if abs(num) <= 1e-300 else float("inf")
    return float(num / den)


def _merge_notes(*notes: str) -> str:
    """multigrid_v4.py:934-936."""
    return "|".join(n for n in notes if n)


def _alignment(ref_norm: float, corr_norm: float, dot_rc: float) -> float:
    """Cosine of the angle between a residual and a correction
    (multigrid_v4.py:855-862)."""
    if ref_norm <= 1e-300 or corr_norm <= 1e-300:
        return 0.0
    return float(dot_rc / (ref_norm * corr_norm))


def lanczos_kappa(op, v, n_iter: int, dot_fn, nrm_fn, axpby_fn, zeros_like):
    """lambda_max / lambda_min of the Lanczos tridiagonal of ``op`` started
    from the unit vector ``v`` (multigrid_v4.py:1083-1115): n_iter steps,
    absolute breakdown at beta < 1e-14, positive Ritz values only; NaN if
    there are none.  The vector arithmetic goes through the callables so the
    recurrence runs on device vectors with the library's own kernels."""
    alpha = np.zeros(n_iter, dtype=np.float64)
    beta = np.zeros(max(n_iter - 1, 0), dtype=np.float64)
    v_prev = zeros_like(v)
    for j in range(n_iter):
        w = op(v)
        a = dot_fn(v, w)
        alpha[j] = a
        axpby_fn(-a, v, 1.0, w)
        if j > 0:
            axpby_fn(-beta[j - 1], v_prev, 1.0, w)
        if j < n_iter - 1:
            b = nrm_fn(w)
            beta[j] = b
            if b < 1e-14:
                alpha = alpha[: j + 1]
                beta = beta[:j]
                break
            v_prev = v
            v = zeros_like(w)
            axpby_fn(1.0 / b, w, 0.0, v)
    T = np.diag(alpha) + np.diag(beta, 1) + np.diag(beta, -1)
    eigs = np.linalg.eigvalsh(T)
    eigs = eigs[eigs > 0]
    if len(eigs) == 0:
        return float("nan")
    return float(eigs[-1] / max(eigs[0], 1e-300))


class GalerkinMatFreeGMG:
    """Production Galerkin GMG hierarchy (multigrid_v4.py:537)."""

    def __init__(self, *, mf_op: MatrixFreeKff, free, free_gpu, nelx: int, nely: int, nelz: int,
                 KE_UNIT, n_levels: int = 4, coarse_smooth_iters: int = 2,
                 fine_smoother: str = "fp32", fine_smoother_degree: int = 2, omega: float = 0.5,
                 fused_op=None, dense_chol_max: int = 5000, coarse_pcg_iters: int = 80,
                 smoother_type: str = "chebyshev", cheb_lower_frac: float = 1.0 / 30.0,
                 level_precisions=None, cycle_type: str = "v",
                 coarse_correction_policy: str = "none",
                 coarse_correction_acceptance_ratio: float = 1.0,
                 coarse_correction_max_scale: float = 1.0, root_enrichment_mode: str = "none",
                 root_enrichment_weight_floor: float = 1e-6,
                 root_enrichment_weight_power: float = 0.5,
                 root_local_correction_mode: str = "none",
                 root_local_correction_max_scale: float = 1.0,
                 root_local_correction_inner_steps: int = 10,
                 root_local_correction_inner_restart=None, inner_krylov_steps: int = 0,
                 inner_krylov_restart=None, coarsest_solve: str = "inverse",
                 graph_apply: bool = True, mgs_sweep=True,
                 coarse_pcg_persistent: bool = True) -> None:
        # argument validation (multigrid_v4.py:600-699), same messages
        if smoother_type not in {"chebyshev", "jacobi"}:
            raise ValueError(f"smoother_type must be 'chebyshev' or 'jacobi'; got {smoother_type!r}")
        if cycle_type not in {"v", "w"}:
            raise ValueError(f"cycle_type must be 'v' or 'w'; got {cycle_type!r}")
        if coarse_correction_policy not in {"none", "residual_line_search"}:
            raise ValueError("coarse_correction_policy must be 'none' or 'residual_line_search'; "
                             f"got {coarse_correction_policy!r}")
        if coarse_correction_acceptance_ratio <= 0.0:
            raise ValueError("coarse_correction_acceptance_ratio must be positive; "
                             f"got {coarse_correction_acceptance_ratio!r}")
        if coarse_correction_max_scale <= 0.0:
            raise ValueError("coarse_correction_max_scale must be positive; "
                             f"got {coarse_correction_max_scale!r}")
        if root_enrichment_mode not in _ENRICH:
            raise ValueError("root_enrichment_mode must be 'none', 'weighted_residual_second', "
                             "'weighted_gap_second', 'void_residual_second', or 'void_gap_second'; "
                             f"got {root_enrichment_mode!r}")
        if not (0.0 < root_enrichment_weight_floor <= 1.0):
            raise ValueError("root_enrichment_weight_floor must lie in (0, 1]; "
                             f"got {root_enrichment_weight_floor!r}")
        if root_enrichment_weight_power <= 0.0:
            raise ValueError("root_enrichment_weight_power must be positive; "
                             f"got {root_enrichment_weight_power!r}")
        if root_local_correction_mode not in _ROOT_LOCAL:
            raise ValueError("root_local_correction_mode must be 'none', 'node_block_line_search', "
                             f"or 'node_block_inner_fgmres'; got {root_local_correction_mode!r}")
        if root_local_correction_max_scale <= 0.0:
            raise ValueError("root_local_correction_max_scale must be positive; "
                             f"got {root_local_correction_max_scale!r}")
        if root_local_correction_inner_steps <= 0:
            raise ValueError("root_local_correction_inner_steps must be positive; "
                             f"got {root_local_correction_inner_steps!r}")
        if root_local_correction_inner_restart is not None and root_local_correction_inner_restart <= 0:
            raise ValueError("root_local_correction_inner_restart must be positive when provided; "
                             f"got {root_local_correction_inner_restart!r}")
        if inner_krylov_steps < 0:
            raise ValueError(f"inner_krylov_steps must be non-negative; got {inner_krylov_steps!r}")
        if inner_krylov_restart is not None and inner_krylov_restart <= 0:
            raise ValueError("inner_krylov_restart must be positive when provided; "
                             f"got {inner_krylov_restart!r}")
        if fine_smoother not in _PREC:
            raise ValueError(f"Unsupported fine smoother '{fine_smoother}'")
        if coarsest_solve not in ("inverse", "triangular"):
            raise ValueError("coarsest_solve must be 'inverse' or 'triangular'")
        if tuple(mf_op.grid) != (int(nelx), int(nely), int(nelz)):
            raise ValueError("nelx/nely/nelz do not match the operator's grid")
        if not np.array_equal(_host(KE_UNIT, np.float64).reshape(24, 24), mf_op._ctx.KE):
            raise ValueError("KE_UNIT differs from the operator's element stiffness")
        self._mf_op = mf_op
        self._free = free
        self._free_gpu = free_gpu
        self._fine_smoother = fine_smoother
        self._fine_degree = max(1, int(fine_smoother_degree))
        self._coarse_smooth_iters = max(1, int(coarse_smooth_iters))
        self._fused_op = fused_op
        self._coarse_correction_policy = coarse_correction_policy
        self._root_enrichment_mode = root_enrichment_mode
        self._root_local_correction_mode = root_local_correction_mode
        self._inner_krylov_steps = int(inner_krylov_steps)
        self._n_levels = max(1, int(n_levels))
        self._level_prec_arg = list(level_precisions) if level_precisions is not None else None

        p = _lib.gf_gmg_params()
        lib = _lib.load()
        lib.gf_gmg_default_params(C.byref(p))
        p.n_levels = self._n_levels
        p.coarse_smooth_iters = int(coarse_smooth_iters)
        p.fine_smoother = _PREC[fine_smoother]
        p.fine_smoother_degree = int(fine_smoother_degree)
        p.omega = float(omega)
        p.dense_chol_max = int(dense_chol_max)
        p.coarse_pcg_iters = int(coarse_pcg_iters)
        p.smoother_type = 0 if smoother_type == "chebyshev" else 1
        p.cheb_lower_frac = float(cheb_lower_frac)
        for i in range(8):
            p.level_prec[i] = -1
        if level_precisions is not None:
            if len(level_precisions) > 8:
                raise ValueError("at most 8 levels are supported")
            for i, name in enumerate(level_precisions):
                if name not in _PREC:
                    raise ValueError(f"unknown level precision {name!r}")
                p.level_prec[i] = _PREC[name]
        p.cycle_type = 0 if cycle_type == "v" else 1
        p.coarse_correction_policy = 0 if coarse_correction_policy == "none" else 1
        p.coarse_correction_acceptance_ratio = float(coarse_correction_acceptance_ratio)
        p.coarse_correction_max_scale = float(coarse_correction_max_scale)
        p.root_enrichment_mode = _ENRICH[root_enrichment_mode]
        p.root_enrichment_weight_floor = float(root_enrichment_weight_floor)
        p.root_enrichment_weight_power = float(root_enrichment_weight_power)
        p.root_local_correction_mode = _ROOT_LOCAL[root_local_correction_mode]
        p.root_local_correction_max_scale = float(root_local_correction_max_scale)
        p.root_local_correction_inner_steps = int(root_local_correction_inner_steps)
        p.root_local_correction_inner_restart = (int(root_local_correction_inner_restart)
                                                 if root_local_correction_inner_restart else 0)
        p.inner_krylov_steps = int(inner_krylov_steps)
        p.inner_krylov_restart = int(inner_krylov_restart) if inner_krylov_restart else 0
        # coarsest dense solve: the inverse L^-T L^-1 formed once per set-up and
        # applied as one GEMV ("inverse"), or the reference's two triangular
        # solves ("triangular"); both from the same Cholesky factor
        p.coarsest_solve = 0 if coarsest_solve == "inverse" else 1
        # record the whole application as a hipGraph after each set-up
        p.graph_apply = 1 if graph_apply else 0
        # outer FGMRES: one cooperative launch per Arnoldi orthogonalisation
        # (True/1 persistent launch, 2 cooperative launch, False/0 per projection)
        if int(mgs_sweep) not in (0, 1, 2):
            raise ValueError("mgs_sweep must be False/0, True/1 or 2")
        p.mgs_sweep = int(mgs_sweep)
        p.coarse_pcg_persistent = 1 if coarse_pcg_persistent else 0
        self._params = p
        h = C.c_void_p()
        _lib.check(lib.gf_gmg_create(C.byref(h), mf_op.ctx_handle, C.byref(p)), "gf_gmg_create")
        self._h = h
        self._lib = lib
        self._E_e_fp64 = None
        self._is_setup = False
        self._info = None
        self._kappa_eff = float("nan")
        info = self._read_info()
        self._n_free = [int(info.n_free[i]) for i in range(self._n_levels)]
        self._dims = [tuple(int(v) for v in info.dims[i]) for i in range(self._n_levels)]

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._lib.gf_gmg_destroy(self._h)
                self._h = None
        except Exception:
            pass

    # ------------------------------------------------------------------ info
    def _read_info(self):
        info = _lib.gf_gmg_info()
        _lib.check(self._lib.gf_gmg_info_get(self._h, C.byref(info)), "gf_gmg_info_get")
        return info

    @property
    def level_stats(self) -> list:
        if not self._is_setup:
            return []
        out = []
        for lv in range(self._n_levels):
            nx, ny, nz = self._dims[lv]
            nnz = 0 if lv == 0 else 243 * (nx + 1) * (ny + 1) * (nz + 1)
            out.append(MultigridLevelStats(
                level=lv,
                kind=(f"fine-matrixfree-{self._fine_smoother}" if lv == 0 else
                      ("level1-galerkin" if lv == 1 else "galerkin-triple-product")),
                n_elem=nx * ny * nz, n_free=self._n_free[lv], nnz=nnz,
                estimated_vram_bytes=8 * (nnz + 3 * self._n_free[lv])))
        return out

    @property
    def omega_used(self) -> float:
        return float(self._info.omega_used) if self._info else 0.5

    @property
    def lambda_max_estimate(self) -> float:
        return float(self._info.lambda_max[0]) if self._info else float("nan")

    @property
    def lambda_max_per_level(self) -> list:
        if not self._info:
            return []
        return [float(self._info.lambda_max[i]) for i in range(self._n_levels)]

    @property
    def level_precisions(self) -> list:
        if self._level_prec_arg is not None:
            lp = list(self._level_prec_arg)
            return lp + ["fp64"] * max(0, self._n_levels - len(lp))
        if self._fine_smoother == "bf16":
            lp = ["bf16"] + (["fp32"] if self._n_levels > 1 else [])
            return lp + ["fp64"] * max(0, self._n_levels - len(lp))
        return [self._fine_smoother] + ["fp64"] * (self._n_levels - 1)

    @property
    def contrast_ratio(self) -> float:
        return float(self._info.contrast_ratio) if self._info else float("nan")

    @property
    def is_high_contrast(self) -> bool:
        return bool(self._info.is_high_contrast) if self._info else False

    @property
    def kappa_eff(self) -> float:
        return self._kappa_eff

    @property
    def coarse_chol_mode(self) -> str:
        if not self._info:
            return "none"
        return {0: "none", 1: "dense", 2: "iterative"}[int(self._info.coarse_chol_mode)]

    @property
    def adaptive_params(self) -> dict:
        p = self._params
        i = self._info
        return {
            "contrast_ratio": self.contrast_ratio,
            "is_high_contrast": self.is_high_contrast,
            "fine_degree_base": self._fine_degree,
            "fine_degree_adapt": int(i.fine_degree_adapt) if i else self._fine_degree,
            "coarse_iters_base": self._coarse_smooth_iters,
            "coarse_iters_adapt": int(i.coarse_iters_adapt) if i else self._coarse_smooth_iters,
            "coarse_correction_policy": self._coarse_correction_policy,
            "coarse_correction_acceptance_ratio": p.coarse_correction_acceptance_ratio,
            "coarse_correction_max_scale": p.coarse_correction_max_scale,
            "root_enrichment_mode": self._root_enrichment_mode,
            "root_enrichment_weight_floor": p.root_enrichment_weight_floor,
            "root_enrichment_weight_power": p.root_enrichment_weight_power,
            "root_local_correction_mode": self._root_local_correction_mode,
            "root_local_correction_max_scale": p.root_local_correction_max_scale,
            "root_local_correction_inner_steps": p.root_local_correction_inner_steps,
            "root_local_correction_inner_restart": (p.root_local_correction_inner_restart or
                                                    max(1, p.root_local_correction_inner_steps)),
            "inner_krylov_steps": p.inner_krylov_steps,
            "inner_krylov_restart": p.inner_krylov_restart or max(1, p.inner_krylov_steps),
        }

    def uses_bf16_smoother(self) -> bool:
        return self._fine_smoother == "bf16"

    def supports_spd_outer(self) -> bool:
        return (self._fine_smoother in {"fp64", "fp32"} and self._coarse_correction_policy == "none"
                and self._root_enrichment_mode == "none"
                and self._root_local_correction_mode == "none" and self._inner_krylov_steps == 0)

    # ----------------------------------------------------------------- setup
    def setup(self, E_e_gpu, fine_diag=None) -> None:
        import torch

        E = _as_dev(E_e_gpu, torch.float64)
        fd = None if fine_diag is None else _as_dev(fine_diag, torch.float64)
        _lib.check(self._lib.gf_gmg_setup(self._h, _lib.ptr(E), _lib.ptr(fd),
                                          _lib.stream_handle()), "gf_gmg_setup")
        self._E_e_fp64 = E
        self._is_setup = True
        self._info = self._read_info()

    def _require_bf16_operator(self) -> None:
        """BF16 smoothing of level 0 runs the fused matvec's bf16 operator
        (multigrid_v4.py:1774-1779): like the reference, a hierarchy built
        without ``fused_op`` refuses to smooth with it."""
        if (self._fine_smoother == "bf16" and self.level_precisions[0] != "fp64"
                and self._fused_op is None):
            raise RuntimeError(
                "BF16 fine smoothing requires enable_fused_cuda=True so the WMMA path exists.")

    # ----------------------------------------------------------------- apply
    def apply(self, rhs):
        import torch

        if not self._is_setup:
            raise RuntimeError("GalerkinMatFreeGMG.setup() must be called before apply().")
        self._require_bf16_operator()
        r = _as_dev(rhs, torch.float64)
        out = torch.empty_like(r)
        _lib.check(self._lib.gf_gmg_apply(self._h, _lib.ptr(r), _lib.ptr(out),
                                          _lib.stream_handle()), "gf_gmg_apply")
        return out

    def vcycle(self, rhs):
        """One base cycle (V or W) without the Krylov / root-local wrappers."""
        import torch

        self._require_bf16_operator()
        r = _as_dev(rhs, torch.float64)
        out = torch.empty_like(r)
        _lib.check(self._lib.gf_gmg_vcycle(self._h, _lib.ptr(r), _lib.ptr(out),
                                           _lib.stream_handle()), "gf_gmg_vcycle")
        return out

    def apply_fine_operator(self, v):
        return self._mf_op.matvec(v, self._E_e_fp64)

    def apply_preconditioned_operator(self, v):
        return self.apply(self.apply_fine_operator(v))

    def probe_quality(self, rhs) -> dict:
        import torch

        z = self.apply(rhs)
        dinv = self.level_dinv(0)
        zj = dinv * _as_dev(rhs, torch.float64)
        zn, jn = nrm2(z), nrm2(zj)
        return {"finite": bool(torch.isfinite(z).all().item()), "pd": dot(_as_dev(rhs, torch.float64), z) > 0.0,
                "z_norm": zn, "jacobi_norm": jn, "z_over_jacobi": zn / max(jn, 1e-30)}

    # ----------------------------------------------------------- diagnostics
    def estimate_kappa_eff(self, n_iter: int = 30, v0=None) -> float:
        """kappa_eff = lambda_max / lambda_min of the Lanczos tridiagonal of the
        applied operator M K (multigrid_v4.py:1057-1117), n_iter steps.  The
        start vector is v0 (normalised) or, by default, a standard normal
        vector from numpy's default_rng(0); the reference draws it from
        CuPy's RandomState(0), which is not available here, so the two start
        from different vectors (the extreme Ritz values, and so the
        estimate, depend on the start vector only weakly)."""
        import torch

        if not self._is_setup:
            raise RuntimeError("Call setup() before estimate_kappa_eff().")
        if (self._coarse_correction_policy != "none" or self._root_enrichment_mode != "none"
                or self._root_local_correction_mode != "none" or self._inner_krylov_steps > 0):
            raise RuntimeError(
                "estimate_kappa_eff requires a linear preconditioner; disable safeguarded "
                "coarse correction, root enrichment, root-local correction, and inner Krylov "
                "wrapping first.")
        n = self._n_free[0]
        if v0 is None:
            v0 = np.random.default_rng(0).standard_normal(n)
        v = _as_dev(v0, torch.float64).clone()
        axpby(1.0 / nrm2(v), v, 0.0, v)
        self._kappa_eff = lanczos_kappa(self.apply_preconditioned_operator, v, int(n_iter), dot,
                                        nrm2, axpby, torch.zeros_like)
        return self._kappa_eff

    def _coarse_decision(self, stv):
        """(raw, applied, note) of a safeguarded coarse correction from its
        device scalars [c.c, r.r, Ac.Ac, r.Ac, applied, degenerate, |trial|]
        (multigrid_v4.py:867-931)."""
        if self._coarse_correction_policy == "none":
            return 1.0, 1.0, ""
        p = self._params
        rr, den, rac, s, degen, tn = stv[1], stv[2], stv[3], stv[4], stv[5], stv[6]
        if degen != 0.0:
            return 0.0, 0.0, "degenerate_correction"
        raw = float(rac / den)
        if tn > p.coarse_correction_acceptance_ratio * math.sqrt(rr):
            return raw, 0.0, "rejected_correction"
        note = ""
        if s <= 1e-12 and raw <= 0.0:
            note = "nonpositive_correction_scale"
        elif abs(s - raw) > 1e-12:
            note = "clipped_correction_scale"
        return raw, float(s), note

    def _root_local_decision(self, sm):
        """(raw, applied, note, |output residual|) of the root-local
        correction from the trace summary (multigrid_v4.py:964-1055)."""
        pre = sm[4]
        mode = self._root_local_correction_mode
        if mode == "none":
            return None, 0.0, "", pre
        cn, rn = math.sqrt(sm[5]), math.sqrt(sm[6])
        if mode == "node_block_inner_fgmres":
            if cn <= 1e-300 or rn <= 1e-300:
                return 0.0, 0.0, "degenerate_root_local_correction", pre
            tn = math.sqrt(sm[7])
            if tn > rn:
                return None, 0.0, "rejected_root_local_inner_fgmres", pre
            return None, 1.0, "root_local_inner_fgmres", tn
        den, rac = sm[7], sm[8]
        if cn <= 1e-300 or rn <= 1e-300 or den <= 1e-300:
            return 0.0, 0.0, "degenerate_root_local_correction", pre
        raw = float(rac / den)
        s = min(float(self._params.root_local_correction_max_scale), max(0.0, raw))
        if s <= 1e-12:
            return raw, 0.0, "nonpositive_root_local_scale", pre
        note = "clipped_root_local_scale" if abs(s - raw) > 1e-12 else ""
        return raw, s, note, sm[9]

    def cycle_trace(self, rhs, *, cycle_type: Optional[str] = None) -> MultigridCycleTrace:
        """One traced V or W cycle on rhs followed by the root-local correction
        (multigrid_v4.py:1361-1728).  The cycle runs in libgfhip.so with the
        kernels of the ordinary cycle, so its output is the iterate that
        apply() returns when inner_krylov_steps = 0; the per-visit norms are
        reduced on the device and read back once."""
        import torch

        if not self._is_setup:
            raise RuntimeError("GalerkinMatFreeGMG.setup() must be called before cycle_trace().")
        self._require_bf16_operator()
        cycle = ("w" if self._params.cycle_type == 1 else "v") if cycle_type is None else cycle_type
        if cycle not in {"v", "w"}:
            raise ValueError(f"cycle_type must be 'v' or 'w'; got {cycle!r}")
        r = _as_dev(rhs, torch.float64)
        out = torch.empty_like(r)
        cap = 1
        while True:
            visits = (_lib.gf_trace_visit * cap)()
            nv = C.c_int32(0)
            sm = (C.c_double * 16)()
            rc = self._lib.gf_gmg_cycle_trace(self._h, _lib.ptr(r), _lib.ptr(out),
                                              0 if cycle == "v" else 1, visits, cap,
                                              C.byref(nv), sm, _lib.stream_handle())
            if rc != 0 and nv.value > cap:
                cap = nv.value
                continue
            _lib.check(rc, "gf_gmg_cycle_trace")
            break
        self._last_trace_output = out
        sm = list(sm)
        chol = self.coarse_chol_mode
        enrich_note = ("root_enrichment:" + self._root_enrichment_mode
                       if self._root_enrichment_mode != "none" else "")
        out_visits = []
        for k in range(nv.value):
            t = visits[k]
            v = list(t.v)
            if t.is_coarsest:
                rn, xn, res = v[0], v[7], v[10]
                out_visits.append(MultigridCycleVisit(
                    visit_id=t.visit_id, level=t.level, cycle_type=cycle, is_coarsest=True,
                    n_coarse_corrections=0, rhs_norm=rn, jacobi_action_norm=v[1],
                    presmooth_solution_norm=0.0, presmooth_residual_norm=rn,
                    presmooth_residual_ratio=1.0, restricted_rhs_norm=0.0,
                    restricted_rhs_ratio=0.0, projector_norm=0.0, coarse_space_gap_norm=0.0,
                    coarse_space_gap_ratio=0.0, coarse_correction_norm=xn,
                    prolongated_correction_norm=xn, correction_alignment=0.0,
                    correction_scale_raw=None, correction_scale_applied=1.0,
                    corrected_residual_norm=res,
                    corrected_residual_vs_presmooth=_safe_ratio(res, rn),
                    postsmooth_solution_norm=xn, postsmooth_residual_norm=res,
                    postsmooth_residual_ratio=_safe_ratio(res, rn), coarsest_mode=chol))
                continue
            rn, prer, pec, cr1 = v[0], v[3], v[8], v[10]
            raw1, app1, note1 = self._coarse_decision(v[11:18])
            kw = {}
            notes2 = ""
            corrected = cr1
            if t.n_coarse_corrections == 2:
                raw2, app2, note2 = self._coarse_decision(v[23:30])
                cr2 = v[22]
                kw = dict(second_restricted_rhs_norm=v[18], second_coarse_correction_norm=v[19],
                          second_prolongated_correction_norm=v[20],
                          second_correction_alignment=_alignment(cr1, v[20], v[21]),
                          second_correction_scale_raw=raw2, second_correction_scale_applied=app2,
                          second_corrected_residual_norm=cr2,
                          second_corrected_residual_vs_corrected=_safe_ratio(cr2, cr1))
                notes2 = note2
                if cycle == "v":
                    corrected = cr2
            out_visits.append(MultigridCycleVisit(
                visit_id=t.visit_id, level=t.level, cycle_type=cycle, is_coarsest=False,
                n_coarse_corrections=int(t.n_coarse_corrections), rhs_norm=rn,
                jacobi_action_norm=v[1], presmooth_solution_norm=v[2],
                presmooth_residual_norm=prer, presmooth_residual_ratio=_safe_ratio(prer, rn),
                restricted_rhs_norm=v[4], restricted_rhs_ratio=_safe_ratio(v[4], prer),
                projector_norm=v[5], coarse_space_gap_norm=v[6],
                coarse_space_gap_ratio=_safe_ratio(v[6], prer), coarse_correction_norm=v[7],
                prolongated_correction_norm=pec, correction_alignment=_alignment(prer, pec, v[9]),
                correction_scale_raw=raw1, correction_scale_applied=app1,
                corrected_residual_norm=corrected,
                corrected_residual_vs_presmooth=_safe_ratio(corrected, prer),
                postsmooth_solution_norm=v[30], postsmooth_residual_norm=v[31],
                postsmooth_residual_ratio=_safe_ratio(v[31], rn),
                notes=_merge_notes(note1, notes2, enrich_note if t.enrichment else "",
                                   "extra_fine_postsmooth" if t.extra_fine_postsmooth else ""),
                **kw))
        raw, applied, lnote, res_norm = self._root_local_decision(sm)
        if lnote:
            root = next((x for x in out_visits if x.level == 0 and not x.is_coarsest), None)
            if root is not None:
                root.notes = _merge_notes(root.notes, lnote)
        rhs_n, jac_n, out_n, rz, pre = sm[0], sm[1], sm[2], sm[3], sm[4]
        ap = self.adaptive_params
        return MultigridCycleTrace(
            cycle_type=cycle, rhs_norm=rhs_n, jacobi_action_norm=jac_n, output_norm=out_n,
            z_over_jacobi=_safe_ratio(out_n, jac_n), rz=rz, pd=bool(rz > 0.0),
            output_residual_norm=res_norm, output_residual_ratio=_safe_ratio(res_norm, rhs_n),
            contrast_ratio=self.contrast_ratio, is_high_contrast=self.is_high_contrast,
            fine_degree_adapt=ap["fine_degree_adapt"], coarse_iters_adapt=ap["coarse_iters_adapt"],
            coarse_correction_policy=self._coarse_correction_policy,
            coarse_correction_acceptance_ratio=ap["coarse_correction_acceptance_ratio"],
            coarse_correction_max_scale=ap["coarse_correction_max_scale"],
            root_enrichment_mode=self._root_enrichment_mode,
            root_enrichment_weight_floor=ap["root_enrichment_weight_floor"],
            root_enrichment_weight_power=ap["root_enrichment_weight_power"],
            root_local_correction_mode=self._root_local_correction_mode,
            root_local_correction_inner_steps=ap["root_local_correction_inner_steps"],
            root_local_correction_inner_restart=ap["root_local_correction_inner_restart"],
            root_local_correction_raw_scale=raw, root_local_correction_applied_scale=applied,
            root_local_correction_residual_ratio=_safe_ratio(res_norm, pre),
            inner_krylov_steps=ap["inner_krylov_steps"],
            inner_krylov_restart=ap["inner_krylov_restart"], visits=out_visits)

    # ------------------------------------------------ level read-outs (tests)
    def level_matvec(self, level: int, x):
        import torch

        x = _as_dev(x, torch.float64)
        y = torch.empty_like(x)
        _lib.check(self._lib.gf_gmg_level_matvec(self._h, int(level), _lib.ptr(x), _lib.ptr(y),
                                                 _lib.stream_handle()), "gf_gmg_level_matvec")
        return y

    def level_dinv(self, level: int):
        import torch

        out = torch.empty(self._n_free[level], dtype=torch.float64, device="cuda")
        _lib.check(self._lib.gf_gmg_level_dinv(self._h, int(level), _lib.ptr(out),
                                               _lib.stream_handle()), "gf_gmg_level_dinv")
        return out

    def level_dense(self, level: int):
        import torch

        n = self._n_free[level]
        out = torch.empty((n, n), dtype=torch.float64, device="cuda")
        _lib.check(self._lib.gf_gmg_level_dense(self._h, int(level), _lib.ptr(out),
                                                _lib.stream_handle()), "gf_gmg_level_dense")
        return out

    def level_free(self, level: int) -> np.ndarray:
        out = np.empty(self._n_free[level], dtype=np.int32)
        _lib.check(self._lib.gf_gmg_level_free(self._h, int(level), out.ctypes.data),
                   "gf_gmg_level_free")
        return out

    def restrict(self, level: int, r):
        import torch

        r = _as_dev(r, torch.float64)
        out = torch.empty(self._n_free[level + 1], dtype=torch.float64, device=r.device)
        _lib.check(self._lib.gf_gmg_restrict(self._h, int(level), _lib.ptr(r), _lib.ptr(out),
                                             _lib.stream_handle()), "gf_gmg_restrict")
        return out

    def prolong(self, level: int, e):
        import torch

        e = _as_dev(e, torch.float64)
        out = torch.empty(self._n_free[level], dtype=torch.float64, device=e.device)
        _lib.check(self._lib.gf_gmg_prolong(self._h, int(level), _lib.ptr(e), _lib.ptr(out),
                                            _lib.stream_handle()), "gf_gmg_prolong")
        return out

    # ------------------------------------------------------- native solve
    def solve(self, b, *, tol=1e-6, maxiter=200, restart=32, history=None, x0=None):
        """FGMRES(restart) with A = K_ff(E of the last setup) and M = apply,
        entirely in libgfhip (one host read-back per outer iteration)."""
        import torch

        if not self._is_setup:
            raise RuntimeError("GalerkinMatFreeGMG.setup() must be called before apply().")
        self._require_bf16_operator()
        bb = _as_dev(b, torch.float64)
        x = x0.clone().contiguous() if x0 is not None else torch.empty_like(bb)
        cap = int(maxiter) + int(maxiter) // max(1, int(restart)) + 2
        hist = np.zeros(cap, dtype=np.float64)
        nh, it, cv = C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(self._lib.gf_gmg_solve(self._h, _lib.ptr(bb), _lib.ptr(x),
                                          1 if x0 is not None else 0, float(tol), int(maxiter),
                                          int(restart), hist.ctypes.data, cap, C.byref(nh),
                                          C.byref(it), C.byref(cv), _lib.stream_handle()),
                   "gf_gmg_solve")
        if history is not None:
            history.extend(float(v) for v in hist[: min(nh.value, cap)])
        return x, int(it.value), bool(cv.value)

    def solve_continue(self, b, *, tol=1e-6, maxiter=200, restart=32, history=None):
        """Continue the last ``solve`` on this hierarchy to ``maxiter`` as if it
        had been called with that limit (gf_gmg_solve_continue).  Valid when
        that solve started from x = 0 with the same ``b`` tensor, ``tol`` and
        ``restart``, converged or stopped at its maxiter inside its first
        restart cycle, and no ``setup`` came since; otherwise raises
        ``ContinueUnavailable``.  ``history`` receives only the new entries."""
        import torch

        if not self._is_setup:
            raise RuntimeError("GalerkinMatFreeGMG.setup() must be called before apply().")
        self._require_bf16_operator()
        bb = _as_dev(b, torch.float64)
        x = torch.empty_like(bb)
        cap = int(maxiter) + int(maxiter) // max(1, int(restart)) + 2
        hist = np.zeros(cap, dtype=np.float64)
        nh, it, cv = C.c_int32(), C.c_int32(), C.c_int32()
        rc = self._lib.gf_gmg_solve_continue(self._h, _lib.ptr(bb), _lib.ptr(x), float(tol),
                                             int(maxiter), int(restart), hist.ctypes.data, cap,
                                             C.byref(nh), C.byref(it), C.byref(cv),
                                             _lib.stream_handle())
        if rc == -4:
            raise ContinueUnavailable(self._lib.gf_last_error().decode())
        _lib.check(rc, "gf_gmg_solve_continue")
        if history is not None:
            history.extend(float(v) for v in hist[: min(nh.value, cap)])
        return x, int(it.value), bool(cv.value)

    def _set_wait_flag(self, which: int) -> None:
        """Raise wait-expiry flag ``which`` (0 sweep, 1 persistent PCG, 2
        one-launch line search) from the host: the next apply / vcycle /
        cycle_trace / solve reports it."""
        _lib.check(self._lib.gf_gmg_set_wait_flag(self._h, int(which), _lib.stream_handle()),
                   "gf_gmg_set_wait_flag")


class ContinueUnavailable(RuntimeError):
    """The last solve of a hierarchy cannot be continued (gf_gmg_solve_continue)."""


def _native_pair(A_op, M_op):
    gmg = getattr(M_op, "__self__", None)
    if not isinstance(gmg, GalerkinMatFreeGMG) or getattr(M_op, "__name__", "") != "apply":
        return None
    if not gmg._is_setup:
        return None
    if isinstance(A_op, FineOperator) and A_op.mf_op is gmg._mf_op:
        E = A_op.E
        if E is gmg._E_e_fp64 or (hasattr(E, "data_ptr") and E.data_ptr() == gmg._E_e_fp64.data_ptr()
                                  and E.dtype == gmg._E_e_fp64.dtype):
            return gmg
    if getattr(A_op, "__self__", None) is gmg and getattr(A_op, "__name__", "") == "apply_fine_operator":
        return gmg
    return None


def _cupy_fgmres(A_op, b, M_op, x0=None, tol: float = 1e-6, maxiter: int = 200,
                 restart: int = 32, history=None):
    """Flexible restarted GMRES, right preconditioning, MGS + Givens
    (multigrid_v4.py:416-534).  Returns (x, total_iters, converged)."""
    import torch

    gmg = _native_pair(A_op, M_op)
    if gmg is not None and restart <= 1024:
        return gmg.solve(b, tol=tol, maxiter=maxiter, restart=restart, history=history, x0=x0)
    b = _as_dev(b, torch.float64)
    x = torch.zeros_like(b) if x0 is None else x0.clone().contiguous()
    b_norm = nrm2(b)
    if b_norm <= 1e-30:
        return x, 0, True
    total = 0
    converged = False
    while total < maxiter:
        r = b.clone()
        axpby(-1.0, _as_dev(A_op(x), torch.float64), 1.0, r)
        beta = nrm2(r)
        if history is not None:
            history.append(beta / b_norm)
        if beta / b_norm <= tol:
            converged = True
            break
        V = [r / beta]
        Z = []
        H = np.zeros((restart + 1, restart))
        cs = np.zeros(restart)
        sn = np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = beta
        inner = min(restart, maxiter - total)
        j_final = 0
        for j in range(inner):
            z = _as_dev(M_op(V[j]), torch.float64)
            Z.append(z)
            w = _as_dev(A_op(z), torch.float64).clone()
            for i in range(j + 1):
                H[i, j] = dot(V[i], w)
                axpby(-H[i, j], V[i], 1.0, w)
            H[j + 1, j] = nrm2(w)
            h_norm = H[j + 1, j]
            if h_norm > 1e-14:
                V.append(w / h_norm)
            for i in range(j):
                tmp = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = tmp
            denom = math.sqrt(H[j, j] ** 2 + H[j + 1, j] ** 2)
            if denom < 1e-300:
                cs[j], sn[j] = 1.0, 0.0
            else:
                cs[j] = H[j, j] / denom
                sn[j] = H[j + 1, j] / denom
            H[j, j] = cs[j] * H[j, j] + sn[j] * H[j + 1, j]
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            total += 1
            j_final = j + 1
            if history is not None:
                history.append(abs(g[j + 1]) / b_norm)
            if abs(g[j + 1]) / b_norm <= tol:
                converged = True
                break
            if h_norm <= 1e-14:
                converged = True
                break
        if j_final > 0:
            R = np.triu(H[:j_final, :j_final])
            rhs = g[:j_final].copy()
            try:
                y = np.linalg.solve(R, rhs)
            except np.linalg.LinAlgError:
                y = np.linalg.lstsq(R, rhs, rcond=None)[0]
            for coeff, zj in zip(y, Z[:j_final]):
                axpby(float(coeff), zj, 1.0, x)
        if converged:
            break
    return x, total, converged


fgmres = _cupy_fgmres
