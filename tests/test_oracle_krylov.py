"""Pins for the oracle's Krylov solvers against textbook constructions.

* ``fgmres`` with a fixed (linear) preconditioner M is right-preconditioned
  GMRES: after k steps its projected residual equals the minimum of
  ||r0 - A M y|| over the Krylov space K_k(A M, r0), computed here from an
  orthonormal basis of that space by a dense least-squares solve; its final
  iterate solves the system (direct solve) and its projected residual equals
  the recomputed true residual.
* restarts: the history restarts with the recomputed residual.
* ``pcg`` converges to the direct solution; successive residuals are
  M-orthogonal (r_i . M r_j = 0) as CG requires.
* ``fixed_pcg_coarse`` with n_iters >= n reproduces the direct solve.
"""

import numpy as np

from oracle import krylov


def _system(n=40, seed=0, spd=False):
    rng = np.random.default_rng(seed)
    if spd:
        Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
        A = Q @ np.diag(np.linspace(1.0, 50.0, n)) @ Q.T
    else:
        A = np.eye(n) * 4.0 + rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    M = np.diag(1.0 / np.abs(np.diag(A))) + 0.01 * rng.standard_normal((n, n)) / n
    return A, b, M


def _min_residual(A, M, r0, k):
    """min_y ||r0 - A M V_k y|| with V_k an orthonormal basis of K_k(AM, r0)."""
    AM = A @ M
    K = np.empty((len(r0), k))
    v = r0 / np.linalg.norm(r0)
    for j in range(k):                     # Arnoldi with double CGS for the basis only
        K[:, j] = v
        w = AM @ v
        for _ in range(2):
            w = w - K[:, : j + 1] @ (K[:, : j + 1].T @ w)
        v = w / np.linalg.norm(w)
    y = np.linalg.lstsq(AM @ K, r0, rcond=None)[0]
    return np.linalg.norm(r0 - AM @ K @ y)


def test_fgmres_is_minimal_residual_over_krylov_space():
    A, b, M = _system()
    h = []
    x, it, conv = krylov.fgmres(lambda v: A @ v, b, lambda v: M @ v, tol=1e-12, maxiter=25,
                                restart=50, history=h)
    bn = np.linalg.norm(b)
    assert abs(h[0] - 1.0) < 1e-15
    for k in range(1, len(h)):
        ref = _min_residual(A, M, b, k) / bn
        assert abs(h[k] - ref) <= 1e-8 * max(ref, 1e-12) + 1e-13, (k, h[k], ref)
    # projected residual equals the recomputed true residual at exit
    assert abs(np.linalg.norm(b - A @ x) / bn - h[-1]) <= 1e-9


def test_fgmres_converges_to_direct_solution_and_flags():
    A, b, M = _system(seed=1)
    x, it, conv = krylov.fgmres(lambda v: A @ v, b, lambda v: M @ v, tol=1e-11, maxiter=200,
                                restart=200)
    assert conv and it <= 40
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)


def test_fgmres_restart_history_and_x0():
    A, b, M = _system(seed=2)
    h = []
    x0 = np.full(len(b), 0.1)
    x, it, conv = krylov.fgmres(lambda v: A @ v, b, lambda v: M @ v, x0=x0, tol=1e-10,
                                maxiter=60, restart=7, history=h)
    bn = np.linalg.norm(b)
    assert abs(h[0] - np.linalg.norm(b - A @ x0) / bn) < 1e-14
    # each cycle of 7 steps starts with a recomputed residual, which equals the
    # projected residual that closed the previous cycle (exact arithmetic)
    assert len(h) == it + (it + 6) // 7 + (0 if not conv else 0)
    for c in range(1, it // 7):
        end_prev = h[c * 8 - 1]
        start = h[c * 8]
        assert abs(start - end_prev) <= 1e-8 * end_prev
    assert conv and np.linalg.norm(b - A @ x) / bn <= 1e-10 * 1.0001


def test_fgmres_zero_rhs_and_lucky_breakdown():
    A, b, M = _system(seed=3)
    x, it, conv = krylov.fgmres(lambda v: A @ v, np.zeros_like(b), lambda v: M @ v)
    assert it == 0 and conv and not np.any(x)
    # exact preconditioner: one step reaches the solution; h_{2,1} ~ 0 ends the cycle
    Ainv = np.linalg.inv(A)
    x, it, conv = krylov.fgmres(lambda v: A @ v, b, lambda v: Ainv @ v, tol=0.0, maxiter=10,
                                restart=10)
    assert it == 1 and conv
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10)


def test_pcg_solution_and_orthogonality():
    A, b, _ = _system(n=30, seed=4, spd=True)
    Dinv = 1.0 / np.diag(A)
    h = []
    x, it, conv = krylov.pcg(A, b, lambda r: Dinv * r, tol=1e-12, maxiter=200, history=h)
    assert conv
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-9)
    # residual history starts at ||b||/||b|| for a zero start
    assert abs(h[0] - 1.0) < 1e-15
    # the three-term recurrence keeps residuals D^-1-orthogonal: rerun and collect
    rs = []
    xk = np.zeros_like(b)
    r = b.copy()
    z = Dinv * r
    p = z.copy()
    for _ in range(6):
        rs.append(r.copy())
        Ap = A @ p
        a = (r @ z) / (p @ Ap)
        xk += a * p
        rn = r - a * Ap
        zn = Dinv * rn
        p = zn + (rn @ zn) / (r @ z) * p
        r, z = rn, zn
    G = np.array([[ri @ (Dinv * rj) for rj in rs] for ri in rs])
    off = G - np.diag(np.diag(G))
    assert np.abs(off).max() <= 1e-8 * np.abs(np.diag(G)).max()
    # the oracle's 6-step iterate equals this recurrence's
    x6, _, _ = krylov.pcg(A, b, lambda r: Dinv * r, tol=0.0, maxiter=7)
    assert np.allclose(x6, xk, rtol=1e-12, atol=1e-14)


def test_pcg_upfront_convergence_check():
    A, b, _ = _system(n=10, seed=5, spd=True)
    x0 = np.linalg.solve(A, b)
    x, it, conv = krylov.pcg(A, b, lambda r: r, x0=x0, tol=1e-6)
    assert it == 0 and conv and np.array_equal(x, x0)


def test_fixed_pcg_coarse_exact_for_enough_iterations():
    A, b, _ = _system(n=25, seed=6, spd=True)
    x = krylov.fixed_pcg_coarse(A, b, 1.0 / np.diag(A), n_iters=60)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)
