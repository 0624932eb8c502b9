"""Oracle: Galerkin geometric multigrid preconditioner (NumPy/SciPy).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Restates multigrid_v4.py:146-1360 and 1730-2024 plus the helpers it imports
from solver_v2.py (62-149, 716-810).  Pinning (``tests/test_oracle_gmg.py``):

* ``scalar_prolongation``       row sums == 1, exact on linear fields (even
                                dims), brute-force weights on tiny grids.
* ``coarse_free_dofs_injection``brute force on tiny grids / face clamp count.
* ``assemble_level1``           equals R K_ff P (assembled fine K) when all
                                fine dims are even; independent per-element
                                P_loc^T KE P_loc restatement for odd dims.
* ``estimate_*_lambda_max``     bounded above by, and converging to, the
                                largest eigenvalue of D^-1 K (scipy eigsh/eigh).
* ``chebyshev_steps``           degree 1 equals x0 + D^-1 r0 / sigma; with a
                                zero-width interval on D^-1 A = lam I it is
                                exact after one step.  The later recurrence
                                steps follow multigrid_v4.py:1804-1810 as
                                written (their first direction is the
                                unscaled z), which is not the textbook
                                polynomial: pinned end to end only.
* ``GalerkinMatFreeGMG``        two-level V-cycle with exact coarsest solve and
                                no line search: linear, symmetric, and B K has
                                real eigenvalues <= 1 with exactly n_coarse of
                                them equal to 1 (Galerkin projector); the
                                safeguarded line search never increases the
                                level residual and takes the clipped minimiser.
  The nonlinear canonical stack as a whole is pinned end-to-end by the
  reference's optimized-density row (C64_MF: 17 iterations, final residual
  within 1e-7 of the table, ``tests/test_oracle_anchor.py``, slow).  Bitwise intermediate values of the
  reference (per-level lambda estimates, accept/reject sequences) cannot be
  pinned without executing the reference: parity unpinned at that
  granularity.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

from .fem import MatrixFreeKff, NodeBlockJacobiMF, edof_hex8
from .krylov import fgmres, fixed_pcg_coarse


# ---------------------------------------------------------------------------
# Transfers (solver_v2.py:62-149, 716-763)
# ---------------------------------------------------------------------------

def scalar_prolongation(nf, nc) -> sp.csr_matrix:
    """Trilinear node prolongation, rows normalised to sum 1.

    nf, nc: element dims (nelx, nely, nelz) of fine and coarse grid.
    Restates solver_v2.py:62-149: fine node i takes parent i/2 (weight 1) if
    even, parents (i-1)/2 and (i+1)/2 (weight 0.5 each) if odd, keeping only
    in-range parents; then each row is divided by its sum.
    """
    fnn = [n + 1 for n in nf]
    cnn = [n + 1 for n in nc]

    def one_d(nfn, ncn):
        rows, cols, w = [], [], []
        for i in range(nfn):
            if i % 2 == 0:
                cand = [(i // 2, 1.0)]
            else:
                cand = [(i // 2, 0.5), (i // 2 + 1, 0.5)]
            for c, wt in cand:
                if 0 <= c < ncn:
                    rows.append(i)
                    cols.append(c)
                    w.append(wt)
        return sp.csr_matrix((w, (rows, cols)), shape=(nfn, ncn))

    Px, Py, Pz = (one_d(a, b) for a, b in zip(fnn, cnn))
    P = sp.kron(sp.kron(Px, Py), Pz, format="csr")
    P.sum_duplicates()
    rs = np.asarray(P.sum(axis=1)).ravel()
    rs = np.where(rs > 1e-12, rs, 1.0)
    return (sp.diags(1.0 / rs, format="csr") @ P).tocsr()


def coarse_free_dofs_injection(nf, nc, free_f) -> np.ndarray:
    """Coarse DOF free iff its injected fine DOF (2i,2j,2k) is free.

    Restates solver_v2.py:716-763 including its linear-index arithmetic
    (fine strides applied to 2*i_c, so a coarse index past a length-1 fine
    axis aliases into the next fine row exactly as in the reference).
    """
    nxf, nyf, nzf = nf
    nxc, nyc, nzc = nc
    ndof_c = 3 * (nxc + 1) * (nyc + 1) * (nzc + 1)
    ndof_f = 3 * (nxf + 1) * (nyf + 1) * (nzf + 1)
    ix, iy, iz = np.meshgrid(np.arange(nxc + 1), np.arange(nyc + 1), np.arange(nzc + 1),
                             indexing="ij")
    fine_node = (2 * ix * (nyf + 1) * (nzf + 1) + 2 * iy * (nzf + 1) + 2 * iz).ravel()
    mask_f = np.zeros(ndof_f, dtype=bool)
    mask_f[free_f] = True
    is_free = np.zeros(ndof_c, dtype=bool)
    ncn = len(fine_node)
    for d in range(3):
        fd = 3 * fine_node + d
        cd = np.arange(ncn) * 3 + d
        ok = fd < ndof_f
        is_free[cd[ok]] = mask_f[fd[ok]]
    return np.where(is_free)[0].astype(np.int32)


def local_prolongation_scalar(s: int, t: int, u: int) -> np.ndarray:
    """8x8 coarse-element -> fine-sub-element trilinear weights
    (multigrid_v4.py:192-216)."""
    pos = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                    (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], dtype=np.float64)
    P = np.zeros((8, 8))
    stu = np.array([s, t, u], dtype=np.float64)
    for fn in range(8):
        cc = (stu + pos[fn]) / 2.0
        for cn in range(8):
            w = 1.0
            for a in range(3):
                w *= (1.0 - cc[a]) if pos[cn, a] == 0.0 else cc[a]
            P[fn, cn] = w
    return P


STU = [(s, t, u) for s in range(2) for t in range(2) for u in range(2)]


def galerkin_q_stu(KE: np.ndarray) -> list:
    """Q_stu = P_24^T KE P_24 for the 8 sub-elements (multigrid_v4.py:261-266)."""
    out = []
    for s, t, u in STU:
        P24 = np.kron(local_prolongation_scalar(s, t, u), np.eye(3))
        out.append(P24.T @ KE @ P24)
    return out


def assemble_level1(nc, nf, free_c, KE, E_fine):
    """Element-wise level-1 Galerkin operator and safe diagonal.

    Restates multigrid_v4.py:219-361: for every coarse element c and every
    (row, col) of its 24x24 pattern with both DOFs free, the raw slot value is
    sum over the 8 sub-elements stu (order s-major) of
    E[fine(c, stu)] * Q_stu[row, col] with fine = min(2c+stu, nel_f-1) per
    axis; duplicate (row, col) slots are then summed.  diag = d where |d|>1e-12
    else 1.
    """
    nxc, nyc, nzc = nc
    nxf, nyf, nzf = nf
    ndof_c = 3 * (nxc + 1) * (nyc + 1) * (nzc + 1)
    n_free_c = len(free_c)
    edof_c = edof_hex8(nxc, nyc, nzc)
    f2f = np.full(ndof_c, -1, dtype=np.int64)
    f2f[free_c] = np.arange(n_free_c)
    rows = np.repeat(edof_c, 24, axis=1).ravel()
    cols = np.tile(edof_c, (1, 24)).ravel()
    keep = (f2f[rows] >= 0) & (f2f[cols] >= 0)
    kidx = np.nonzero(keep)[0]
    elem = kidx // 576
    lke = kidx % 576
    cx = elem // (nyc * nzc)
    cy = (elem // nzc) % nyc
    cz = elem % nzc
    Q = galerkin_q_stu(KE)
    raw = np.zeros(len(kidx))
    for k, (s, t, u) in enumerate(STU):
        fx = np.minimum(2 * cx + s, nxf - 1)
        fy = np.minimum(2 * cy + t, nyf - 1)
        fz = np.minimum(2 * cz + u, nzf - 1)
        fe = fx * nyf * nzf + fy * nzf + fz
        raw = raw + E_fine[fe] * Q[k].ravel()[lke]
    K = sp.coo_matrix((raw, (f2f[rows[kidx]], f2f[cols[kidx]])),
                      shape=(n_free_c, n_free_c)).tocsr()
    K.sum_duplicates()
    K.sort_indices()
    d = K.diagonal()
    d = np.where(np.abs(d) > 1e-12, d, 1.0)
    return K, d


# ---------------------------------------------------------------------------
# Spectral estimates (multigrid_v4.py:364-405)
# ---------------------------------------------------------------------------

def estimate_fine_lambda_max(mf_op, E, d_inv, n_iter=20) -> float:
    d = 1.0 / d_inv
    v = np.ones(d_inv.shape[0])
    v /= np.sqrt(np.dot(v * d, v))
    lam = 1.0
    for _ in range(n_iter):
        Kv = mf_op.matvec(v, E)
        vDv = float(np.dot(v * d, v))
        vKv = float(np.dot(v, Kv))
        lam = vKv / max(vDv, 1e-300)
        Av = d_inv * Kv
        nrm = float(np.sqrt(np.dot(Av * d, Av)))
        if nrm <= 1e-300:
            break
        v = Av / nrm
    return lam


def estimate_sparse_lambda_max(K, d_inv, n_iter=10) -> float:
    v = np.ones(d_inv.shape[0])
    v /= float(np.linalg.norm(v))
    lam = 1.0
    for _ in range(n_iter):
        Kv = K @ v
        Av = d_inv * Kv
        nrm = float(np.linalg.norm(Av))
        if nrm <= 1e-300:
            break
        lam = float(np.dot(v, Kv)) / max(float(np.dot(v, v / d_inv)), 1e-300)
        v = Av / nrm
    return max(lam, 1e-6)


def chebyshev_steps(A_op, rhs, x0, d_inv, lam_max, lower_frac, degree):
    """Saad Alg. 12.1 form used by the reference (multigrid_v4.py:1783-1812)."""
    lam_min = lower_frac * lam_max
    sigma = (lam_max + lam_min) * 0.5
    delta = (lam_max - lam_min) * 0.5
    x = x0
    r = rhs - A_op(x)
    z = d_inv * r
    alpha = 2.0 / (lam_max + lam_min)
    d_vec = z
    x = x + alpha * d_vec
    for _ in range(1, degree):
        r = rhs - A_op(x)
        z = d_inv * r
        alpha_new = 1.0 / (sigma - (delta ** 2) * alpha * 0.25)
        d_vec = alpha_new * (z + (delta ** 2) * alpha * 0.25 * d_vec)
        x = x + d_vec
        alpha = alpha_new
    return x


def lanczos_kappa(op, v, n_iter):
    """lambda_max / lambda_min of the n_iter-step Lanczos tridiagonal of op
    from the unit vector v (multigrid_v4.py:1083-1115): absolute breakdown at
    beta < 1e-14, positive Ritz values only, NaN if there are none."""
    alpha = np.zeros(n_iter)
    beta = np.zeros(max(n_iter - 1, 0))
    v_prev = np.zeros_like(v)
    for j in range(n_iter):
        w = op(v)
        alpha[j] = float(np.dot(v, w))
        w = w - alpha[j] * v
        if j > 0:
            w = w - beta[j - 1] * v_prev
        if j < n_iter - 1:
            beta[j] = float(np.linalg.norm(w))
            if beta[j] < 1e-14:
                alpha, beta = alpha[: j + 1], beta[:j]
                break
            v_prev, v = v, w / beta[j]
    ev = np.linalg.eigvalsh(np.diag(alpha) + np.diag(beta, 1) + np.diag(beta, -1))
    ev = ev[ev > 0]
    return float("nan") if len(ev) == 0 else float(ev[-1] / max(ev[0], 1e-300))


def _safe_ratio(num, den):
    """multigrid_v4.py:840-843."""
    if abs(den) <= 1e-300:
        return 0.0 if abs(num) <= 1e-300 else float("inf")
    return float(num / den)


def _merge_notes(*notes):
    return "|".join(n for n in notes if n)


def _alignment(ref, corr):
    """multigrid_v4.py:855-862."""
    a, b = float(np.linalg.norm(ref)), float(np.linalg.norm(corr))
    if a <= 1e-300 or b <= 1e-300:
        return 0.0
    return float(np.dot(ref, corr) / (a * b))


@dataclass
class Decision:
    x: object
    residual: object
    raw_scale: Optional[float]
    applied_scale: float
    note: str = ""


class GalerkinMatFreeGMG:
    """Restatement of multigrid_v4.py:537-2024 (production path)."""

    def __init__(self, *, mf_op: MatrixFreeKff, free, nelx, nely, nelz, KE_UNIT,
                 n_levels=4, coarse_smooth_iters=2, fine_smoother="fp32",
                 fine_smoother_degree=2, omega=0.5, dense_chol_max=5000,
                 coarse_pcg_iters=80, smoother_type="chebyshev", cheb_lower_frac=1.0 / 30.0,
                 level_precisions=None, cycle_type="v", coarse_correction_policy="none",
                 coarse_correction_acceptance_ratio=1.0, coarse_correction_max_scale=1.0,
                 root_enrichment_mode="none", root_enrichment_weight_floor=1e-6,
                 root_enrichment_weight_power=0.5, root_local_correction_mode="none",
                 root_local_correction_max_scale=1.0, root_local_correction_inner_steps=10,
                 root_local_correction_inner_restart=None, inner_krylov_steps=0,
                 inner_krylov_restart=None):
        assert smoother_type in ("chebyshev", "jacobi")
        assert cycle_type in ("v", "w")
        assert coarse_correction_policy in ("none", "residual_line_search")
        assert root_local_correction_mode in ("none", "node_block_line_search",
                                              "node_block_inner_fgmres")
        self._mf = mf_op
        self._KE = np.asarray(KE_UNIT, dtype=np.float64)
        self._fine_smoother = fine_smoother
        self._fine_degree = max(1, int(fine_smoother_degree))
        self._coarse_iters = max(1, int(coarse_smooth_iters))
        self._omega_max = float(omega)
        self._omega_used = float(omega)
        self._dense_chol_max = int(dense_chol_max)
        self._coarse_pcg_iters = int(coarse_pcg_iters)
        self._n_levels = max(1, int(n_levels))
        self._smoother_type = smoother_type
        self._cheb_lower_frac = float(cheb_lower_frac)
        self._level_prec = list(level_precisions) if level_precisions is not None else None
        self._cycle_type = cycle_type
        self._cc_policy = coarse_correction_policy
        self._cc_ratio = float(coarse_correction_acceptance_ratio)
        self._cc_max = float(coarse_correction_max_scale)
        self._re_mode = root_enrichment_mode
        self._re_floor = float(root_enrichment_weight_floor)
        self._re_power = float(root_enrichment_weight_power)
        self._rl_mode = root_local_correction_mode
        self._rl_max = float(root_local_correction_max_scale)
        self._rl_steps = int(root_local_correction_inner_steps)
        self._rl_restart = (int(root_local_correction_inner_restart)
                            if root_local_correction_inner_restart is not None
                            else max(1, self._rl_steps))
        self._ik_steps = int(inner_krylov_steps)
        self._ik_restart = (int(inner_krylov_restart) if inner_krylov_restart is not None
                            else max(1, self._ik_steps))
        self._lambda_max_est = float("nan")
        self._lambda_per_level = []
        self._lam_E_max = float("nan")
        self._contrast_ratio = float("nan")
        self._void_frac = float("nan")
        self._is_high_contrast = False
        self._coarse_iters_adapt = coarse_smooth_iters
        self._fine_degree_adapt = self._fine_degree
        self._coarse_chol_mode = "none"
        self._chol_L = None
        self._is_setup = False
        self._root_bj = None
        self._re_weights = None

        dims = [(nelx, nely, nelz)]
        for _ in range(1, self._n_levels):
            nx, ny, nz = dims[-1]
            dims.append((max(1, nx // 2), max(1, ny // 2), max(1, nz // 2)))
        self._dims = dims
        self._free = [np.asarray(free)]
        self._n_free = [len(free)]
        self._P = []
        for lv in range(self._n_levels - 1):
            fc = coarse_free_dofs_injection(dims[lv], dims[lv + 1], self._free[lv])
            self._free.append(fc)
            self._n_free.append(len(fc))
            Psc = scalar_prolongation(dims[lv], dims[lv + 1])
            Pv = sp.kron(Psc, sp.eye(3, format="csr"), format="csr")
            self._P.append(Pv[self._free[lv], :][:, fc].tocsr())
        self._R = [P.T.tocsr() for P in self._P]
        self._K = [None] * self._n_levels
        self._d_inv = [None] * self._n_levels
        self._K32 = [None] * self._n_levels
        self._d_inv32 = [None] * self._n_levels
        self._E64 = None
        self._E32 = None

    # -- properties (multigrid_v4.py:771-836) -------------------------------
    @property
    def omega_used(self):
        return self._omega_used

    @property
    def lambda_max_estimate(self):
        return self._lambda_max_est

    @property
    def lambda_max_per_level(self):
        return list(self._lambda_per_level)

    @property
    def contrast_ratio(self):
        return self._contrast_ratio

    @property
    def is_high_contrast(self):
        return self._is_high_contrast

    @property
    def adaptive_params(self):
        return {"contrast_ratio": self._contrast_ratio,
                "is_high_contrast": self._is_high_contrast,
                "fine_degree_adapt": self._fine_degree_adapt,
                "coarse_iters_adapt": self._coarse_iters_adapt}

    @property
    def level_precisions(self):
        return list(self._level_prec) if self._level_prec else []

    # -- setup (multigrid_v4.py:1119-1341) ----------------------------------
    def setup(self, E, fine_diag=None):
        self._E64 = np.asarray(E, dtype=np.float64)
        self._E32 = self._E64.astype(np.float32)
        if self._level_prec is None:
            if self._fine_smoother == "bf16":
                self._level_prec = ["bf16"] + (["fp32"] if self._n_levels > 1 else [])
                self._level_prec += ["fp64"] * max(0, self._n_levels - len(self._level_prec))
            else:
                self._level_prec = [self._fine_smoother] + ["fp64"] * (self._n_levels - 1)
        while len(self._level_prec) < self._n_levels:
            self._level_prec.append("fp64")
        d0 = fine_diag if fine_diag is not None else self._mf.extract_diagonal(self._E64)
        d0 = np.where(np.abs(d0) > 1e-12, d0, 1.0)
        self._d_inv[0] = 1.0 / d0
        if self._re_mode != "none":
            wb = np.abs(d0) / max(float(np.max(np.abs(d0))), 1e-300)
            wb = np.clip(wb, self._re_floor, 1.0)
            if abs(self._re_power - 1.0) > 1e-12:
                wb = wb ** self._re_power
            self._re_weights = wb
        else:
            self._re_weights = None
        if self._rl_mode != "none":
            if self._root_bj is None:
                m = self._mf
                self._root_bj = NodeBlockJacobiMF(m._edof, m._KE, m._free, m._n_free, m._ndof)
            self._root_bj.update(self._E64, scalar_diag=d0)
        else:
            self._root_bj = None
        E_max = float(np.max(self._E64))
        prev = self._lam_E_max
        reest = np.isnan(prev) or abs(E_max - prev) / max(abs(prev), 1e-12) > 0.10
        if reest:
            self._lambda_max_est = estimate_fine_lambda_max(self._mf, self._E64, self._d_inv[0])
            self._lam_E_max = E_max
        self._omega_used = min(self._omega_max, 0.9 * 2.0 / max(self._lambda_max_est, 1e-6))
        if len(self._lambda_per_level) != self._n_levels:
            self._lambda_per_level = [float("nan")] * self._n_levels
        self._lambda_per_level[0] = self._lambda_max_est
        self._K32 = [None] * self._n_levels
        self._d_inv32 = [None] * self._n_levels

        E_min = float(np.min(self._E64))
        self._contrast_ratio = E_max / max(E_min, 1e-300)
        n_solid = int(np.count_nonzero(self._E64 > float(E_max * 1e-8)))
        void_frac = 1.0 - n_solid / max(len(self._E64), 1)
        self._void_frac = void_frac
        bc, bf = self._coarse_iters, self._fine_degree
        if self._contrast_ratio > 1e5 or void_frac > 0.8:
            self._coarse_iters_adapt, self._fine_degree_adapt, self._is_high_contrast = max(bc, 6), max(bf, 4), True
        elif self._contrast_ratio > 100 or void_frac > 0.5:
            self._coarse_iters_adapt, self._fine_degree_adapt, self._is_high_contrast = max(bc, 4), max(bf, 3), True
        elif self._contrast_ratio > 10:
            self._coarse_iters_adapt, self._fine_degree_adapt, self._is_high_contrast = max(bc, 3), bf, False
        else:
            self._coarse_iters_adapt, self._fine_degree_adapt, self._is_high_contrast = bc, bf, False

        if self._n_levels > 1:
            K1, d1 = assemble_level1(self._dims[1], self._dims[0], self._free[1], self._KE, self._E64)
            self._K[1] = K1
            self._d_inv[1] = 1.0 / d1
            if reest or np.isnan(self._lambda_per_level[1]):
                self._lambda_per_level[1] = estimate_sparse_lambda_max(K1, self._d_inv[1])
            if self._level_prec[1] == "fp32":
                self._K32[1] = K1.astype(np.float32)
                self._d_inv32[1] = self._d_inv[1].astype(np.float32)
        for lv in range(2, self._n_levels):
            Kc = (self._R[lv - 1] @ (self._K[lv - 1] @ self._P[lv - 1])).tocsr()
            Kc.sum_duplicates()
            Kc.sort_indices()
            dc = Kc.diagonal()
            dc = np.where(np.abs(dc) > 1e-12, dc, 1.0)
            self._K[lv] = Kc
            self._d_inv[lv] = 1.0 / dc
            if reest or np.isnan(self._lambda_per_level[lv]):
                self._lambda_per_level[lv] = estimate_sparse_lambda_max(Kc, self._d_inv[lv])
            if self._level_prec[lv] == "fp32":
                self._K32[lv] = Kc.astype(np.float32)
                self._d_inv32[lv] = self._d_inv[lv].astype(np.float32)

        if self._n_levels == 1:
            self._coarse_chol_mode, self._chol_L = "none", None
        else:
            Kc = self._K[self._n_levels - 1]
            n = Kc.shape[0]
            if n <= self._dense_chol_max:
                Kd = Kc.toarray()
                try:
                    L = np.linalg.cholesky(Kd)
                    if not np.all(np.isfinite(L)):
                        raise RuntimeError("non-finite L")
                    self._chol_L, self._coarse_chol_mode = L, "dense"
                except Exception:
                    eps = max(float(Kd.diagonal().mean()) * 1e-8, 1e-14)
                    try:
                        L = np.linalg.cholesky(Kd + eps * np.eye(n))
                        if not np.all(np.isfinite(L)):
                            raise RuntimeError("non-finite")
                        self._chol_L, self._coarse_chol_mode = L, "dense"
                    except Exception:
                        self._chol_L, self._coarse_chol_mode = None, "iterative"
            else:
                self._chol_L, self._coarse_chol_mode = None, "iterative"
        self._is_setup = True

    # -- operators ------------------------------------------------------------
    def apply_fine_operator(self, v):
        return self._mf.matvec(v, self._E64)

    def _level_matvec(self, lv, v):
        return self.apply_fine_operator(v) if lv == 0 else self._K[lv] @ v

    def _fine_matvec_low(self, v32):
        # multigrid_v4.py:1765-1780
        if self._fine_smoother == "fp32":
            return self._mf.matvec(v32, self._E32)
        if self._fine_smoother == "bf16":
            return self._mf.matvec_bf16(v32, self._E32)
        raise ValueError(f"Unsupported fine smoother '{self._fine_smoother}'")

    # -- corrections (multigrid_v4.py:867-1055) -------------------------------
    def _apply_coarse_correction(self, lv, x, residual, corr):
        if self._cc_policy == "none":
            return Decision(x + corr, residual - self._level_matvec(lv, corr), 1.0, 1.0)
        cn = float(np.linalg.norm(corr))
        rn = float(np.linalg.norm(residual))
        if cn <= 1e-300 or rn <= 1e-300:
            return Decision(x, residual, 0.0, 0.0, "degenerate_correction")
        Ac = self._level_matvec(lv, corr)
        denom = float(np.dot(Ac, Ac))
        if denom <= 1e-300:
            return Decision(x, residual, 0.0, 0.0, "degenerate_correction")
        raw = float(np.dot(residual, Ac) / denom)
        s = min(self._cc_max, max(0.0, raw))
        trial = residual - s * Ac
        tn = float(np.linalg.norm(trial))
        if tn > self._cc_ratio * rn:
            return Decision(x, residual, raw, 0.0, "rejected_correction")
        note = ""
        if s <= 1e-12 and raw <= 0.0:
            note = "nonpositive_correction_scale"
        elif abs(s - raw) > 1e-12:
            note = "clipped_correction_scale"
        return Decision(x + s * corr, trial, raw, s, note)

    def _root_enrichment_rhs(self, residual):
        w = self._re_weights
        vw = np.clip(1.0 - w, self._re_floor, 1.0)
        R0, P0 = self._R[0], self._P[0]
        if self._re_mode == "weighted_residual_second":
            return R0 @ (w * residual)
        if self._re_mode == "weighted_gap_second":
            return R0 @ (w * (residual - P0 @ (R0 @ residual)))
        if self._re_mode == "void_residual_second":
            return R0 @ (vw * residual)
        if self._re_mode == "void_gap_second":
            return R0 @ (vw * (residual - P0 @ (R0 @ residual)))
        raise RuntimeError(self._re_mode)

    def _apply_root_local_correction(self, x, residual):
        if self._rl_mode == "none":
            return Decision(x, residual, None, 0.0)
        if self._rl_mode == "node_block_inner_fgmres":
            corr, _, _ = fgmres(self.apply_fine_operator, residual, self._root_bj.apply,
                                tol=0.0, maxiter=self._rl_steps, restart=self._rl_restart)
            cn = float(np.linalg.norm(corr))
            rn = float(np.linalg.norm(residual))
            if cn <= 1e-300 or rn <= 1e-300:
                return Decision(x, residual, 0.0, 0.0, "degenerate_root_local_correction")
            trial = residual - self.apply_fine_operator(corr)
            tn = float(np.linalg.norm(trial))
            if tn > rn:
                return Decision(x, residual, None, 0.0, "rejected_root_local_inner_fgmres")
            return Decision(x + corr, trial, None, 1.0, "root_local_inner_fgmres")
        corr = self._root_bj.apply(residual)
        cn = float(np.linalg.norm(corr))
        rn = float(np.linalg.norm(residual))
        if cn <= 1e-300 or rn <= 1e-300:
            return Decision(x, residual, 0.0, 0.0, "degenerate_root_local_correction")
        Ac = self.apply_fine_operator(corr)
        denom = float(np.dot(Ac, Ac))
        if denom <= 1e-300:
            return Decision(x, residual, 0.0, 0.0, "degenerate_root_local_correction")
        raw = float(np.dot(residual, Ac) / denom)
        s = min(self._rl_max, max(0.0, raw))
        if s <= 1e-12:
            return Decision(x, residual, raw, 0.0, "nonpositive_root_local_scale")
        note = "clipped_root_local_scale" if abs(s - raw) > 1e-12 else ""
        return Decision(x + s * corr, residual - s * Ac, raw, s, note)

    # -- apply (multigrid_v4.py:1730-1767) ------------------------------------
    def _apply_base_cycle(self, rhs):
        return self._wcycle(0, rhs) if self._cycle_type == "w" else self._vcycle(0, rhs)

    def apply(self, rhs):
        if not self._is_setup:
            raise RuntimeError("GalerkinMatFreeGMG.setup() must be called before apply().")
        rhs = np.asarray(rhs, dtype=np.float64)
        if self._ik_steps > 0:
            x, _, _ = fgmres(self.apply_fine_operator, rhs, self._apply_base_cycle,
                             tol=0.0, maxiter=self._ik_steps, restart=self._ik_restart)
        else:
            x = self._apply_base_cycle(rhs)
        if self._rl_mode == "none":
            return x
        residual = rhs - self.apply_fine_operator(x)
        return self._apply_root_local_correction(x, residual).x

    def probe_quality(self, rhs):
        z = self.apply(rhs)
        zj = self._d_inv[0] * rhs
        return {"finite": bool(np.all(np.isfinite(z))), "pd": float(np.dot(rhs, z)) > 0.0,
                "z_norm": float(np.linalg.norm(z)), "jacobi_norm": float(np.linalg.norm(zj)),
                "z_over_jacobi": float(np.linalg.norm(z) / max(float(np.linalg.norm(zj)), 1e-30))}

    # -- smoothers (multigrid_v4.py:1814-1906) --------------------------------
    def _lower_frac(self):
        lf = self._cheb_lower_frac
        if self._is_high_contrast:
            lf = min(self._cheb_lower_frac, 1.0 / 100.0)
        return lf

    def _smooth_fine(self, x, rhs):
        prec = self._level_prec[0]
        lm = self._lambda_max_est
        if self._smoother_type == "chebyshev":
            if prec == "fp64":
                return chebyshev_steps(self.apply_fine_operator, rhs, x.astype(np.float64, copy=True),
                                       self._d_inv[0], lm, self._lower_frac(), self._fine_degree_adapt)
            x32 = x.astype(np.float32, copy=True)
            r32 = rhs.astype(np.float32)
            d32 = self._d_inv[0].astype(np.float32)
            return chebyshev_steps(self._fine_matvec_low, r32, x32, d32, np.float32(lm),
                                   np.float32(self._lower_frac()),
                                   self._fine_degree_adapt).astype(np.float64)
        if prec == "fp64":
            xc = x.astype(np.float64, copy=True)
            for _ in range(self._fine_degree_adapt):
                xc = xc + self._omega_used * self._d_inv[0] * (rhs - self.apply_fine_operator(xc))
            return xc
        x32 = x.astype(np.float32, copy=True)
        r32 = rhs.astype(np.float32)
        d32 = self._d_inv[0].astype(np.float32)
        for _ in range(self._fine_degree_adapt):
            x32 = x32 + self._omega_used * d32 * (r32 - self._fine_matvec_low(x32))
        return x32.astype(np.float64)

    def _smooth_coarse(self, lv, x, rhs):
        prec = self._level_prec[lv] if lv < len(self._level_prec) else "fp64"
        lm = self._lambda_per_level[lv] if lv < len(self._lambda_per_level) else 1.0
        lm = lm if np.isfinite(lm) else 1.0
        lf = self._lower_frac()
        it = self._coarse_iters_adapt
        if prec == "fp32" and self._K32[lv] is not None:
            K32, d32 = self._K32[lv], self._d_inv32[lv]
            x32 = x.astype(np.float32, copy=True)
            r32 = rhs.astype(np.float32)
            if self._smoother_type == "chebyshev":
                x32 = chebyshev_steps(lambda v: K32 @ v, r32, x32, d32, np.float32(lm),
                                      np.float32(lf), it)
            else:
                for _ in range(it):
                    x32 = x32 + self._omega_used * d32 * (r32 - K32 @ x32)
            return x32.astype(np.float64)
        K, d = self._K[lv], self._d_inv[lv]
        if self._smoother_type == "chebyshev":
            return chebyshev_steps(lambda v: K @ v, rhs, x, d, lm, lf, it)
        xc = x
        for _ in range(it):
            xc = xc + self._omega_used * d * (rhs - K @ xc)
        return xc

    def _solve_coarsest(self, rhs):
        if self._n_levels == 1:
            return self._smooth_fine(np.zeros_like(rhs), rhs)
        lv = self._n_levels - 1
        if self._coarse_chol_mode == "dense":
            y = sla.solve_triangular(self._chol_L, rhs, lower=True)
            return sla.solve_triangular(self._chol_L, y, lower=True, trans="T")
        it = self._coarse_pcg_iters
        if self._is_high_contrast:
            it = max(self._coarse_pcg_iters, 200)
        return fixed_pcg_coarse(self._K[lv], rhs, self._d_inv[lv], n_iters=it)

    # -- cycles (multigrid_v4.py:1929-2024) -----------------------------------
    def _vcycle(self, lv, rhs):
        if lv == self._n_levels - 1:
            return self._solve_coarsest(rhs)
        if lv == 0:
            x = self._smooth_fine(np.zeros_like(rhs), rhs)
            residual = rhs - self.apply_fine_operator(x)
        else:
            x = self._smooth_coarse(lv, np.zeros_like(rhs), rhs)
            residual = rhs - self._K[lv] @ x
        e_c = self._vcycle(lv + 1, self._R[lv] @ residual)
        Pe = self._P[lv] @ e_c
        if self._cc_policy == "none":
            x = x + Pe
            corrected = residual - self._level_matvec(lv, Pe)
        else:
            dec = self._apply_coarse_correction(lv, x, residual, Pe)
            x, corrected = dec.x, dec.residual
        if lv == 0 and self._re_mode != "none":
            e2 = self._vcycle(lv + 1, self._root_enrichment_rhs(corrected))
            Pe2 = self._P[lv] @ e2
            if self._cc_policy == "none":
                x = x + Pe2
            else:
                x = self._apply_coarse_correction(lv, x, corrected, Pe2).x
        x = self._smooth_fine(x, rhs) if lv == 0 else self._smooth_coarse(lv, x, rhs)
        if lv == 0 and self._is_high_contrast:
            x = self._smooth_fine(x, rhs)
        return x

    def _wcycle(self, lv, rhs):
        if lv == self._n_levels - 1:
            return self._solve_coarsest(rhs)
        if lv == 0:
            x = self._smooth_fine(np.zeros_like(rhs), rhs)
            residual = rhs - self.apply_fine_operator(x)
        else:
            x = self._smooth_coarse(lv, np.zeros_like(rhs), rhs)
            residual = rhs - self._K[lv] @ x
        e_c = self._wcycle(lv + 1, self._R[lv] @ residual)
        Pe = self._P[lv] @ e_c
        if self._cc_policy == "none":
            x = x + Pe
            r2 = rhs - (self.apply_fine_operator(x) if lv == 0 else self._K[lv] @ x)
        else:
            dec = self._apply_coarse_correction(lv, x, residual, Pe)
            x, r2 = dec.x, dec.residual
        rc2 = (self._root_enrichment_rhs(r2) if (lv == 0 and self._re_mode != "none")
               else self._R[lv] @ r2)
        e2 = self._wcycle(lv + 1, rc2)
        Pe2 = self._P[lv] @ e2
        if self._cc_policy == "none":
            x = x + Pe2
        else:
            x = self._apply_coarse_correction(lv, x, r2, Pe2).x
        x = self._smooth_fine(x, rhs) if lv == 0 else self._smooth_coarse(lv, x, rhs)
        if lv == 0 and self._is_high_contrast:
            x = self._smooth_fine(x, rhs)
        return x

    # -- diagnostics (multigrid_v4.py:1057-1117, 1361-1728) -------------------
    def estimate_kappa_eff(self, n_iter=30, v0=None):
        """Lanczos on M K (multigrid_v4.py:1057-1117).  Start vector: v0, or
        numpy default_rng(0) (the reference's CuPy RandomState(0) stream is
        not available)."""
        if not self._is_setup:
            raise RuntimeError("Call setup() before estimate_kappa_eff().")
        if (self._cc_policy != "none" or self._re_mode != "none" or self._rl_mode != "none"
                or self._ik_steps > 0):
            raise RuntimeError(
                "estimate_kappa_eff requires a linear preconditioner; disable safeguarded "
                "coarse correction, root enrichment, root-local correction, and inner Krylov "
                "wrapping first.")
        n = len(self._free[0])
        v = np.random.default_rng(0).standard_normal(n) if v0 is None else np.array(v0, float)
        v = v / float(np.linalg.norm(v))
        return lanczos_kappa(lambda u: self.apply(self.apply_fine_operator(u)), v, int(n_iter))

    def cycle_trace(self, rhs, cycle_type=None):
        """One traced cycle plus the root-local correction, as plain dicts with
        the fields of the reference's MultigridCycleTrace / MultigridCycleVisit
        (multigrid_v4.py:1361-1728)."""
        if not self._is_setup:
            raise RuntimeError("GalerkinMatFreeGMG.setup() must be called before cycle_trace().")
        rhs = np.asarray(rhs, dtype=np.float64)
        cycle = self._cycle_type if cycle_type is None else cycle_type
        if cycle not in {"v", "w"}:
            raise ValueError(f"cycle_type must be 'v' or 'w'; got {cycle!r}")
        visits, counter = [], [0]
        out = self._trace_level(0, rhs, visits, counter, cycle)
        pre = rhs - self.apply_fine_operator(out)
        dec = self._apply_root_local_correction(out, pre)
        out = dec.x
        nrm = lambda u: float(np.linalg.norm(u))
        rn, jn, on, resn = nrm(rhs), nrm(self._d_inv[0] * rhs), nrm(out), nrm(dec.residual)
        rz = float(np.dot(rhs, out))
        visits.sort(key=lambda t: t["visit_id"])
        if dec.note:
            root = next((t for t in visits if t["level"] == 0 and not t["is_coarsest"]), None)
            if root is not None:
                root["notes"] = _merge_notes(root["notes"], dec.note)
        ap = self.adaptive_params
        return dict(
            cycle_type=cycle, rhs_norm=rn, jacobi_action_norm=jn, output_norm=on,
            z_over_jacobi=_safe_ratio(on, jn), rz=rz, pd=bool(rz > 0.0),
            output_residual_norm=resn, output_residual_ratio=_safe_ratio(resn, rn),
            contrast_ratio=self._contrast_ratio, is_high_contrast=self._is_high_contrast,
            fine_degree_adapt=self._fine_degree_adapt, coarse_iters_adapt=self._coarse_iters_adapt,
            coarse_correction_policy=self._cc_policy,
            coarse_correction_acceptance_ratio=self._cc_ratio,
            coarse_correction_max_scale=self._cc_max, root_enrichment_mode=self._re_mode,
            root_enrichment_weight_floor=self._re_floor,
            root_enrichment_weight_power=self._re_power,
            root_local_correction_mode=self._rl_mode,
            root_local_correction_inner_steps=self._rl_steps,
            root_local_correction_inner_restart=self._rl_restart,
            root_local_correction_raw_scale=dec.raw_scale,
            root_local_correction_applied_scale=dec.applied_scale,
            root_local_correction_residual_ratio=_safe_ratio(resn, nrm(pre)),
            inner_krylov_steps=self._ik_steps, inner_krylov_restart=self._ik_restart,
            visits=visits, output=out)

    def _trace_level(self, lv, rhs, visits, counter, cycle):
        """_trace_vcycle_level / _trace_wcycle_level (multigrid_v4.py:1438-1728)."""
        nrm = lambda u: float(np.linalg.norm(u))
        vid = counter[0]
        counter[0] += 1
        rn = nrm(rhs)
        jn = nrm(self._d_inv[lv] * rhs)
        if lv == self._n_levels - 1:
            x = self._solve_coarsest(rhs)
            res = nrm(rhs - self._level_matvec(lv, x))
            xn = nrm(x)
            visits.append(dict(
                visit_id=vid, level=lv, cycle_type=cycle, is_coarsest=True, n_coarse_corrections=0,
                rhs_norm=rn, jacobi_action_norm=jn, presmooth_solution_norm=0.0,
                presmooth_residual_norm=rn, presmooth_residual_ratio=1.0, restricted_rhs_norm=0.0,
                restricted_rhs_ratio=0.0, projector_norm=0.0, coarse_space_gap_norm=0.0,
                coarse_space_gap_ratio=0.0, coarse_correction_norm=xn,
                prolongated_correction_norm=xn, correction_alignment=0.0,
                corrected_residual_norm=res, corrected_residual_vs_presmooth=_safe_ratio(res, rn),
                correction_scale_raw=None, correction_scale_applied=1.0,
                second_restricted_rhs_norm=None, second_coarse_correction_norm=None,
                second_prolongated_correction_norm=None, second_correction_alignment=None,
                second_correction_scale_raw=None, second_correction_scale_applied=None,
                second_corrected_residual_norm=None, second_corrected_residual_vs_corrected=None,
                postsmooth_solution_norm=xn, postsmooth_residual_norm=res,
                postsmooth_residual_ratio=_safe_ratio(res, rn), coarsest_mode=self._coarse_chol_mode,
                notes=""))
            return x
        z = np.zeros_like(rhs)
        xs = self._smooth_fine(z, rhs) if lv == 0 else self._smooth_coarse(lv, z, rhs)
        r = rhs - self._level_matvec(lv, xs)
        prn = nrm(r)
        rc = self._R[lv] @ r
        proj = self._P[lv] @ rc
        ec = self._trace_level(lv + 1, rc, visits, counter, cycle)
        Pe = self._P[lv] @ ec
        d1 = self._apply_coarse_correction(lv, xs, r, Pe)
        x, r1 = d1.x, d1.residual
        c1n = nrm(r1)
        second = {}
        d2 = None
        enote = ""
        corrected_n = c1n
        if cycle == "w" or (lv == 0 and self._re_mode != "none"):
            if lv == 0 and self._re_mode != "none":
                rc2 = self._root_enrichment_rhs(r1)
                enote = "root_enrichment:" + self._re_mode
            else:
                rc2 = self._R[lv] @ r1
            e2 = self._trace_level(lv + 1, rc2, visits, counter, cycle)
            Pe2 = self._P[lv] @ e2
            d2 = self._apply_coarse_correction(lv, x, r1, Pe2)
            x = d2.x
            c2n = nrm(d2.residual)
            if cycle == "v":
                corrected_n = c2n
            second = dict(second_restricted_rhs_norm=nrm(rc2), second_coarse_correction_norm=nrm(e2),
                          second_prolongated_correction_norm=nrm(Pe2),
                          second_correction_alignment=_alignment(r1, Pe2),
                          second_correction_scale_raw=d2.raw_scale,
                          second_correction_scale_applied=d2.applied_scale,
                          second_corrected_residual_norm=c2n,
                          second_corrected_residual_vs_corrected=_safe_ratio(c2n, c1n))
        else:
            second = dict(second_restricted_rhs_norm=None, second_coarse_correction_norm=None,
                          second_prolongated_correction_norm=None, second_correction_alignment=None,
                          second_correction_scale_raw=None, second_correction_scale_applied=None,
                          second_corrected_residual_norm=None,
                          second_corrected_residual_vs_corrected=None)
        extra = ""
        if lv == 0:
            x = self._smooth_fine(x, rhs)
            if self._is_high_contrast:
                x = self._smooth_fine(x, rhs)
                extra = "extra_fine_postsmooth"
        else:
            x = self._smooth_coarse(lv, x, rhs)
        post = nrm(rhs - self._level_matvec(lv, x))
        visits.append(dict(
            visit_id=vid, level=lv, cycle_type=cycle, is_coarsest=False,
            n_coarse_corrections=2 if d2 is not None else 1, rhs_norm=rn, jacobi_action_norm=jn,
            presmooth_solution_norm=nrm(xs), presmooth_residual_norm=prn,
            presmooth_residual_ratio=_safe_ratio(prn, rn), restricted_rhs_norm=nrm(rc),
            restricted_rhs_ratio=_safe_ratio(nrm(rc), prn), projector_norm=nrm(proj),
            coarse_space_gap_norm=nrm(r - proj), coarse_space_gap_ratio=_safe_ratio(nrm(r - proj), prn),
            coarse_correction_norm=nrm(ec), prolongated_correction_norm=nrm(Pe),
            correction_alignment=_alignment(r, Pe), corrected_residual_norm=corrected_n,
            corrected_residual_vs_presmooth=_safe_ratio(corrected_n, prn),
            correction_scale_raw=d1.raw_scale, correction_scale_applied=d1.applied_scale,
            postsmooth_solution_norm=nrm(x), postsmooth_residual_norm=post,
            postsmooth_residual_ratio=_safe_ratio(post, rn), coarsest_mode="",
            notes=_merge_notes(d1.note, d2.note if d2 is not None else "", enote, extra), **second))
        return x
