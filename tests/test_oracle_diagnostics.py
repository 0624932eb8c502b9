"""Pins for the oracle's diagnostics (multigrid_v4.py:1057-1117, 1361-1728).

* Lanczos recurrence (``lanczos_kappa``): on a diagonal operator with n <= the
  step count it breaks down after n steps and returns max/min of the spectrum
  exactly; with fewer steps the Ritz values lie inside the spectrum, so the
  estimate is below the true ratio and grows with the step count.
* ``estimate_kappa_eff`` applies M K (the preconditioner after the operator):
  it equals the recurrence on the assembled B K from the same start vector,
  and differs from the one on K B; it refuses nonlinear stacks.
* ``cycle_trace``: its output is the iterate of ``apply`` (V cycles, where the
  traced and production cycles do the same arithmetic); visit counts are
  L (V), 2^L - 1 (W) and 1 + 2(L-1) (V with root enrichment); the first
  visit's post-smoothing residual is |rhs - K x| recomputed; an exact dense
  coarsest solve leaves a residual at rounding level; a line-search scale is
  inside [0, max_scale] and equals the raw scale unless a note says why.
"""

import numpy as np
import pytest

from oracle import fem, gmg


def test_lanczos_exact_on_small_diagonal():
    lam = np.array([0.5, 1.0, 2.0, 3.0, 7.5, 10.0])
    v = np.random.default_rng(1).standard_normal(len(lam))
    v /= np.linalg.norm(v)
    k = gmg.lanczos_kappa(lambda u: lam * u, v, 30)
    assert abs(k - 20.0) <= 1e-10 * 20.0


def test_lanczos_partial_is_inside_the_spectrum_and_grows():
    lam = np.linspace(1.0, 100.0, 200)
    v = np.random.default_rng(2).standard_normal(len(lam))
    v /= np.linalg.norm(v)
    ks = [gmg.lanczos_kappa(lambda u: lam * u, v, m) for m in (3, 6, 12, 24)]
    assert all(1.0 < k <= 100.0 * (1 + 1e-12) for k in ks)
    assert all(b >= a for a, b in zip(ks, ks[1:]))


def _two_level(dims=(8, 4, 4), **kw):
    spec = fem.cantilever_spec(*dims)
    bc, free, edof, F, mf = fem.build_components(spec)
    g = gmg.GalerkinMatFreeGMG(mf_op=mf, free=free, nelx=dims[0], nely=dims[1], nelz=dims[2],
                               KE_UNIT=fem.KE_UNIT_3D, n_levels=2, fine_smoother="fp64",
                               fine_smoother_degree=3, **kw)
    rng = np.random.default_rng(3)
    E = np.where(rng.random(edof.shape[0]) < 0.5, 1.0, 1e-3)
    g.setup(E)
    K = fem.assemble_K(edof, E, fem.KE_UNIT_3D, bc.ndof)[free][:, free].toarray()
    return g, K, len(free)


def test_kappa_eff_is_lanczos_on_MK():
    g, K, n = _two_level()
    B = np.column_stack([g.apply(e) for e in np.eye(n)])
    v0 = np.random.default_rng(0).standard_normal(n)
    v = v0 / np.linalg.norm(v0)
    k = g.estimate_kappa_eff(n_iter=20)
    k_mk = gmg.lanczos_kappa(lambda u: B @ (K @ u), v, 20)
    k_km = gmg.lanczos_kappa(lambda u: K @ (B @ u), v, 20)
    assert abs(k - k_mk) <= 1e-8 * k_mk
    assert abs(k - k_km) > 1e-6 * k_mk
    g2, _, _ = _two_level(coarse_correction_policy="residual_line_search")
    with pytest.raises(RuntimeError):
        g2.estimate_kappa_eff()


def _setup(dims, n_levels=3, **kw):
    spec = fem.cantilever_spec(*dims)
    bc, free, edof, F, mf = fem.build_components(spec)
    g = gmg.GalerkinMatFreeGMG(mf_op=mf, free=free, nelx=dims[0], nely=dims[1], nelz=dims[2],
                               KE_UNIT=fem.KE_UNIT_3D, n_levels=n_levels, fine_smoother="fp64",
                               **kw)
    rng = np.random.default_rng(7)
    E = np.where(rng.random(edof.shape[0]) < 0.4, 1.0, 1e-3)
    g.setup(E)
    return g, mf, E, len(free)


@pytest.mark.parametrize("kw,cycle,nvis", [
    (dict(), "v", 3),
    (dict(coarse_correction_policy="residual_line_search"), "v", 3),
    (dict(coarse_correction_policy="residual_line_search", root_enrichment_mode="weighted_gap_second"),
     "v", 5),
    (dict(coarse_correction_policy="residual_line_search",
          root_local_correction_mode="node_block_line_search"), "v", 3),
    (dict(root_local_correction_mode="node_block_inner_fgmres"), "v", 3),
    (dict(coarse_correction_policy="residual_line_search", cycle_type="w"), "w", 7),
])
def test_cycle_trace_structure_and_output(kw, cycle, nvis):
    g, mf, E, n = _setup((12, 6, 6), **kw)
    rhs = np.random.default_rng(9).standard_normal(n)
    tr = g.cycle_trace(rhs)
    vis = tr["visits"]
    assert [v["visit_id"] for v in vis] == list(range(nvis))
    assert tr["cycle_type"] == cycle
    assert np.array_equal(tr["output"], g.apply(rhs))
    assert abs(tr["rhs_norm"] - np.linalg.norm(rhs)) <= 1e-14 * tr["rhs_norm"]
    assert tr["pd"] == (float(rhs @ tr["output"]) > 0.0)
    coarse = [v for v in vis if v["is_coarsest"]]
    assert all(v["level"] == 2 for v in coarse)
    assert len(coarse) == {3: 1, 5: 2, 7: 4}[nvis]
    for v in coarse:
        assert v["corrected_residual_norm"] <= 1e-9 * v["rhs_norm"]
    if kw.get("root_local_correction_mode", "none") == "none" and cycle == "v":
        x = g._apply_base_cycle(rhs)
        r = np.linalg.norm(rhs - mf.matvec(x, E))
        assert abs(vis[0]["postsmooth_residual_norm"] - r) <= 1e-12 * r
    for v in vis:
        if v["is_coarsest"]:
            continue
        assert -1.0 - 1e-12 <= v["correction_alignment"] <= 1.0 + 1e-12
        if kw.get("coarse_correction_policy") == "residual_line_search":
            s, raw = v["correction_scale_applied"], v["correction_scale_raw"]
            assert 0.0 <= s <= 1.0
            if "correction" not in v["notes"]:
                assert abs(s - raw) <= 1e-12
        else:
            assert v["correction_scale_raw"] == 1.0 and v["correction_scale_applied"] == 1.0
        assert (v["second_corrected_residual_norm"] is None) == (v["n_coarse_corrections"] == 1)
