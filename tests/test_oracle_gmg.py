"""Pins for the oracle's multigrid pieces.

* trilinear prolongation: rows sum to 1 and linear fields are reproduced
  exactly (even dims); an odd trailing fine node keeps its single parent.
* injection free sets: x=0 face clamp gives exactly the coarse nodes with
  ix > 0; the reference's aliasing arithmetic on a length-1 axis.
* level-1 element-wise Galerkin = P^T K_ff P with K_ff assembled by scipy
  (all fine dims even), including partly clamped boundaries.
* power-iteration estimates are Rayleigh quotients of D^-1 K (never above its
  largest eigenvalue) and converge to it.
* Chebyshev smoother: degree 1 is the closed form x0 + D^-1 r0 / sigma; a zero
  interval width (lower_frac = 1) on D^-1 A = lam I solves in one step.
* a two-level V-cycle with an exact coarsest solve, no line search and a
  symmetric smoother is a fixed, symmetric linear operator B, and B K has
  real eigenvalues <= 1 with exactly n_coarse of them equal to 1 (exact
  coarse-space solve); the safeguarded line search never raises the residual.
"""

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fem, gmg


def _linear(nodes, coef):
    ix, iy, iz = np.meshgrid(*(np.arange(n) for n in nodes), indexing="ij")
    return (coef[0] + coef[1] * ix + coef[2] * iy + coef[3] * iz).ravel()


@pytest.mark.parametrize("nf,nc", [((4, 2, 6), (2, 1, 3)), ((8, 4, 4), (4, 2, 2))])
def test_prolongation_rows_and_linear_exactness(nf, nc):
    P = gmg.scalar_prolongation(nf, nc)
    assert np.allclose(np.asarray(P.sum(axis=1)).ravel(), 1.0, atol=1e-15)
    coef = (0.3, 1.7, -0.4, 2.2)
    fine_nodes = [n + 1 for n in nf]
    coarse_nodes = [n + 1 for n in nc]
    fc = _linear(coarse_nodes, [coef[0], 2 * coef[1], 2 * coef[2], 2 * coef[3]])
    assert np.allclose(P @ fc, _linear(fine_nodes, coef), atol=1e-13)


def test_prolongation_odd_trailing_node_single_parent():
    # 3 fine elements (nodes 0..3) onto 1 coarse element (nodes 0..1):
    # node 3 is odd with parents 1 and 2; parent 2 is out of range -> weight 1 on 1
    P = gmg.scalar_prolongation((3, 2, 2), (1, 1, 1)).toarray()
    Px = P.reshape(4, 3, 3, 2, 2, 2)[:, 0, 0, :, 0, 0]
    assert np.allclose(Px, [[1, 0], [0.5, 0.5], [0, 1], [0, 1]])


def test_injection_free_set_face_clamp():
    spec = fem.cantilever_spec(8, 4, 6)
    bc = fem.generate_bc(spec)
    fc = gmg.coarse_free_dofs_injection((8, 4, 6), (4, 2, 3), bc.free_dofs)
    nodes = fc // 3
    ix = nodes // (3 * 4)
    assert np.all(ix > 0)
    assert len(fc) == 3 * 4 * 3 * 4


def test_injection_aliasing_on_unit_axis():
    # nelz = 1 fine -> coarse nelz = max(1, 1 // 2) = 1: coarse node z=1 maps to
    # fine linear index 2 * iz, which lies in the next fine (iy) row.
    nf, nc = (2, 2, 1), (1, 1, 1)
    ndof_f = 3 * 3 * 3 * 2
    free_f = np.arange(ndof_f)
    free_f = free_f[(free_f // 3) != 3]          # fix fine node 3 = (0, 1, 1)
    fc = gmg.coarse_free_dofs_injection(nf, nc, free_f)
    # coarse node (0,0,1) -> fine linear 2 = (0, 1, 0): free; coarse node
    # (0,1,0) -> fine 2*(nz_f+1) = 4 = (0, 2, 0)... ; coarse (0,0,1) never hits 3
    coarse_nodes = set((fc // 3).tolist())
    # brute force with the reference's arithmetic: fine = 2ix*(3*2) + 2iy*2 + 2iz
    expect = set()
    for ix in range(2):
        for iy in range(2):
            for iz in range(2):
                f = 2 * ix * 3 * 2 + 2 * iy * 2 + 2 * iz
                if f < 18 and f != 3:
                    expect.add((ix * 2 + iy) * 2 + iz)
    assert coarse_nodes == expect


@pytest.mark.parametrize("dims,kind", [((8, 4, 4), "tip"), ((6, 4, 2), "bridge")])
def test_level1_galerkin_equals_ptkp(dims, kind):
    spec = (fem.cantilever_spec(*dims) if kind == "tip"
            else fem._bridge(dims, "bridge_small", (0.0, 0.25, 0.5)))
    bc, free, edof, F, mf = fem.build_components(spec)
    nf = (spec.nelx, spec.nely, spec.nelz)
    nc = tuple(max(1, n // 2) for n in nf)
    if any(n % 2 for n in nf):
        pytest.skip("odd fine dims: the reference's clamp makes K1 differ from P^T K P")
    rng = np.random.default_rng(1)
    E = np.where(rng.random(edof.shape[0]) < 0.5, 1.0, 1e-3)
    fc = gmg.coarse_free_dofs_injection(nf, nc, free)
    K1, d1 = gmg.assemble_level1(nc, nf, fc, fem.KE_UNIT_3D, E)
    Kf = fem.assemble_K(edof, E, fem.KE_UNIT_3D, 3 * np.prod([n + 1 for n in nf]))
    Kff = Kf[free][:, free]
    Psc = gmg.scalar_prolongation(nf, nc)
    P = sp.kron(Psc, sp.eye(3), format="csr")[free][:, fc]
    ref = (P.T @ Kff @ P).toarray()
    assert np.abs(K1.toarray() - ref).max() <= 1e-13 * np.abs(ref).max()
    d = np.diag(ref)
    assert np.array_equal(d1, np.where(np.abs(d) > 1e-12, K1.diagonal(), 1.0))


def test_spectral_estimates_are_rayleigh_quotients():
    spec = fem.cantilever_spec(6, 3, 3)
    bc, free, edof, F, mf = fem.build_components(spec)
    E = np.where(np.random.default_rng(2).random(edof.shape[0]) < 0.6, 1.0, 1e-2)
    d = mf.extract_diagonal(E)
    d_inv = 1.0 / d
    K = fem.assemble_K(edof, E, fem.KE_UNIT_3D, bc.ndof)[free][:, free].toarray()
    lam = np.linalg.eigvals(np.diag(d_inv) @ K).real.max()
    est20 = gmg.estimate_fine_lambda_max(mf, E, d_inv, n_iter=20)
    assert est20 <= lam * (1 + 1e-12)
    est = gmg.estimate_fine_lambda_max(mf, E, d_inv, n_iter=3000)
    assert abs(est - lam) <= 1e-6 * lam
    Ks = sp.csr_matrix(K)
    e10 = gmg.estimate_sparse_lambda_max(Ks, d_inv, n_iter=10)
    assert e10 <= lam * (1 + 1e-12)
    e = gmg.estimate_sparse_lambda_max(Ks, d_inv, n_iter=3000)
    assert abs(e - lam) <= 1e-6 * lam


def test_chebyshev_closed_forms():
    rng = np.random.default_rng(3)
    n = 12
    a = rng.uniform(1.0, 5.0, n)
    A = np.diag(a)
    b = rng.standard_normal(n)
    x0 = rng.standard_normal(n)
    d_inv = 1.0 / a * 0.7
    lam = 3.0
    x1 = gmg.chebyshev_steps(lambda v: A @ v, b, x0, d_inv, lam, 1 / 30, 1)
    sigma = 0.5 * (lam + lam / 30)
    assert np.allclose(x1, x0 + d_inv * (b - A @ x0) / sigma, rtol=1e-15, atol=1e-15)
    # D^-1 A = lam I and a zero-width interval: exact after the first step
    d_inv = lam / a
    for deg in (1, 2, 5):
        x = gmg.chebyshev_steps(lambda v: A @ v, b, x0, d_inv, lam, 1.0, deg)
        assert np.allclose(x, b / a, rtol=1e-13)


def _two_level(dims=(8, 4, 4), **kw):
    spec = fem.cantilever_spec(*dims)
    bc, free, edof, F, mf = fem.build_components(spec)
    g = gmg.GalerkinMatFreeGMG(mf_op=mf, free=free, nelx=dims[0], nely=dims[1], nelz=dims[2],
                               KE_UNIT=fem.KE_UNIT_3D, n_levels=2, fine_smoother="fp64",
                               fine_smoother_degree=3, **kw)
    E = np.ones(edof.shape[0])
    g.setup(E)
    K = fem.assemble_K(edof, E, fem.KE_UNIT_3D, bc.ndof)[free][:, free]
    return g, K, len(free)


def test_two_level_cycle_is_symmetric_linear_and_contracting():
    g, K, n = _two_level()
    assert g._coarse_chol_mode == "dense" and not g.is_high_contrast
    rng = np.random.default_rng(4)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    Bx, By = g.apply(x), g.apply(y)
    assert np.allclose(g.apply(2.0 * x - 3.0 * y), 2.0 * Bx - 3.0 * By, rtol=1e-11, atol=1e-12)
    assert abs(x @ By - y @ Bx) <= 1e-11 * abs(x @ By)
    # I - B K = S (I - P Kc^-1 P^T K) S with S the (K-self-adjoint) smoother
    # polynomial and an exact Galerkin coarse solve: the middle factor is a
    # K-orthogonal projector, so B K has real eigenvalues <= 1, and exactly
    # n_coarse of them equal 1 (the coarse space is solved exactly).  A wrong
    # restriction, a non-Galerkin coarse operator or unpaired smoothing breaks it.
    Bd = np.column_stack([g.apply(e) for e in np.eye(n)])
    w = np.linalg.eigvals(Bd @ K.toarray())
    assert np.abs(w.imag).max() <= 1e-8
    assert w.real.max() <= 1.0 + 1e-10
    assert int(np.sum(np.abs(w.real - 1.0) <= 1e-8)) == g._n_free[1]


def test_line_search_never_increases_residual():
    g, K, n = _two_level(coarse_correction_policy="residual_line_search")
    rng = np.random.default_rng(5)
    for _ in range(10):
        r = rng.standard_normal(n)
        corr = rng.standard_normal(n)
        dec = g._apply_coarse_correction(0, np.zeros(n), r, corr)
        assert np.linalg.norm(dec.residual) <= np.linalg.norm(r) * (1 + 1e-14)
        # the accepted scale minimises ||r - s K c|| over [0, 1]
        if dec.applied_scale > 0:
            Kc = K @ corr
            s = np.clip((r @ Kc) / (Kc @ Kc), 0.0, 1.0)
            assert abs(dec.applied_scale - s) <= 1e-12 * max(s, 1e-30)
