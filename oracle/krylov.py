"""Oracle: Krylov solvers (NumPy).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Pinning (``tests/test_oracle_krylov.py``): ``fgmres`` with a fixed linear
preconditioner against a textbook right-preconditioned GMRES built from a dense
least-squares problem on the same Krylov space, and its solution against a
direct solve; ``pcg`` against a direct solve and against the CG three-term
identities; ``fixed_pcg_coarse`` against the exact solve on SPD systems where
n_iters >= n.
"""

from __future__ import annotations

import numpy as np


def fgmres(A_op, b, M_op, x0=None, tol=1e-6, maxiter=200, restart=32, history=None):
    """Flexible restarted GMRES, right preconditioning, MGS + Givens.

    Restates multigrid_v4.py:416-534 step for step: restart loop with
    r = b - A x; history of beta/||b|| then |g_{j+1}|/||b|| per step;
    projected-residual stopping; lucky breakdown at an absolute 1e-14;
    back-solve of the triangular factor with a least-squares fallback;
    x += sum_j y_j z_j.  Returns (x, total_iters, converged).
    """
    x = np.zeros_like(b) if x0 is None else x0.copy()
    b_norm = float(np.linalg.norm(b))
    if b_norm <= 1e-30:
        return x, 0, True
    total = 0
    converged = False
    while total < maxiter:
        r = b - A_op(x)
        beta = float(np.linalg.norm(r))
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
            z = M_op(V[j])
            Z.append(z)
            w = A_op(z)
            for i in range(j + 1):
                H[i, j] = float(np.dot(V[i], w))
                w = w - H[i, j] * V[i]
            H[j + 1, j] = float(np.linalg.norm(w))
            h_norm = H[j + 1, j]
            if h_norm > 1e-14:
                V.append(w / h_norm)
            for i in range(j):
                tmp = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = tmp
            denom = np.sqrt(H[j, j] ** 2 + H[j + 1, j] ** 2)
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
            for c, z in zip(y, Z[:j_final]):
                x = x + float(c) * z
        if converged:
            break
    return x, total, converged


def pcg(A, b, M_op, x0=None, tol=1e-5, maxiter=1000, history=None):
    """PCG with the up-front residual check (solver_v2.py:551-623)."""
    Aop = A if callable(A) else (lambda v: A @ v)
    x = x0.copy() if x0 is not None else np.zeros_like(b)
    r = b - Aop(x)
    bb = float(np.dot(b, b))
    tol_sq = tol ** 2 * bb
    rr0 = float(np.dot(r, r))
    if history is not None:
        history.append((rr0 / bb) ** 0.5 if bb > 0 else 0.0)
    if rr0 <= tol_sq:
        return x, 0, True
    z = M_op(r)
    p = z.copy()
    rz = float(np.dot(r, z))
    converged = False
    iters = 1
    for _ in range(maxiter - 1):
        Ap = Aop(p)
        pAp = float(np.dot(p, Ap))
        alpha = rz / pAp
        x = x + alpha * p
        r = r - alpha * Ap
        rr = float(np.dot(r, r))
        if history is not None:
            history.append((rr / bb) ** 0.5 if bb > 0 else 0.0)
        if rr <= tol_sq:
            converged = True
            break
        z = M_op(r)
        iters += 1
        rz_new = float(np.dot(r, z))
        beta = rz_new / rz
        p = z + beta * p
        rz = rz_new
    return x, iters, converged


def fixed_pcg_coarse(K, b, d_inv, n_iters=40):
    """Fixed-count Jacobi PCG, zero start, no stopping (solver_v2.py:781-810)."""
    x = np.zeros_like(b)
    r = b.copy()
    z = d_inv * r
    p = z.copy()
    rz = r.dot(z)
    for _ in range(n_iters):
        Ap = K @ p
        pAp = p.dot(Ap)
        alpha = rz / (pAp + 1e-300)
        x = x + alpha * p
        r = r - alpha * Ap
        z = d_inv * r
        rz_n = r.dot(z)
        beta = rz_n / (rz + 1e-300)
        p = z + beta * p
        rz = rz_n
    return x
