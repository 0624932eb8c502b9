"""Probe continuation and the wait-expiry error path, through the C ABI.

gf_gmg_solve_continue: the kept floor's solve continues the probe's Arnoldi
cycle.  The reference re-solves from scratch
(run_gmg_floor_detector_prospective.py:263-277 then 313-351); continuing must
give that solve bit for bit: the iterate, the iteration count and the whole
residual history.

gf_gmg_set_wait_flag: a grid-wide wait that gave up (forced from the host,
no kernel is made to wait) is reported by the next apply / vcycle / solve,
names the wait, and leaves the hierarchy usable with unchanged results.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STACK = dict(n_levels=4, fine_smoother="fp64", coarse_correction_policy="residual_line_search",
             root_local_correction_mode="node_block_inner_fgmres",
             root_local_correction_inner_steps=20, root_local_correction_inner_restart=20,
             inner_krylov_steps=10, inner_krylov_restart=10)


def _runner(dims, **kw):
    from gpu_fem import policy as gpol

    st = dict(STACK)
    st.update(kw.pop("stack", {}))
    return gpol.GuardedFloorPolicy.from_cantilever(dims, load="tip", penal=4.5, **kw, **st)


def _mask(seed, dims, p):
    from oracle import policy as opol

    return opol.bernoulli_mask(seed, dims, p).ravel()


@pytest.mark.parametrize("dims,seed,p,rho_min,first", [((24, 12, 12), 43, 0.35, 1e-12, 10),
                                                        ((24, 12, 12), 5, 0.2, 1e-3, 25),
                                                        ((16, 8, 8), 43, 0.35, 1e-12, 100)])
def test_solve_continue_equals_longer_solve(dims, seed, p, rho_min, first):
    import torch
    from gpu_fem.multigrid_v4 import ContinueUnavailable

    r = _runner(dims)
    E = torch.as_tensor(r.moduli(rho_min, _mask(seed, dims, p)), device="cuda")
    r.gmg.setup(E)
    h_full: list = []
    x_full, it_full, cv_full = r.gmg.solve(r.F, tol=1e-6, maxiter=300, restart=300,
                                           history=h_full)
    r.gmg.setup(E)
    h1: list = []
    _, it1, cv1 = r.gmg.solve(r.F, tol=1e-6, maxiter=first, restart=300, history=h1)
    h2: list = []
    x2, it2, cv2 = r.gmg.solve_continue(r.F, tol=1e-6, maxiter=300, restart=300, history=h2)
    torch.cuda.synchronize()
    print(f"{dims} seed {seed} p {p} rho {rho_min:g}: first {it1} conv {cv1}, "
          f"continued {it2} vs {it_full}")
    assert (it2, cv2) == (it_full, cv_full)
    assert h1 + h2 == h_full
    assert torch.equal(x2, x_full)
    if cv1:
        assert it1 == it_full
    # a second continuation of a finished (converged or maxiter) solve
    if cv2:
        x3, it3, _ = r.gmg.solve_continue(r.F, tol=1e-6, maxiter=300, restart=300)
        assert it3 == it2 and torch.equal(x3, x2)
    # no continuation after a set-up, from a given x0, or to a smaller maxiter
    r.gmg.setup(E)
    with pytest.raises(ContinueUnavailable):
        r.gmg.solve_continue(r.F, tol=1e-6, maxiter=300, restart=300)
    r.gmg.solve(r.F, tol=1e-6, maxiter=first, restart=300, x0=torch.zeros_like(r.F))
    with pytest.raises(ContinueUnavailable):
        r.gmg.solve_continue(r.F, tol=1e-6, maxiter=300, restart=300)
    r.gmg.solve(r.F, tol=1e-6, maxiter=first, restart=300)
    with pytest.raises(ContinueUnavailable):
        r.gmg.solve_continue(r.F, tol=1e-6, maxiter=300, restart=200)


@pytest.mark.parametrize("seed,p", [(43, 0.35), (41, 0.08), (7, 0.2), (47, 0.3)])
def test_policy_probe_continuation_rows_identical(seed, p):
    """Whole guarded-policy rows with and without the continuation: equal
    decisions, iterations, histories and bitwise iterates.  probe_iters 20
    (r50 at 10) so that kept floors continue mid-cycle."""
    import torch

    dims = (24, 12, 12)
    kw = dict(probe_iters=20, probe_r50=10)
    mask = _mask(seed, dims, p)
    rows = []
    for reuse in (False, True):
        r = _runner(dims, reuse_probe=reuse, **kw)
        rows.append(r.run_state(mask, seed=seed, probability=p))
    a, b = rows
    torch.cuda.synchronize()
    print(f"seed {seed} p {p}: trigger {a['trigger']} attempted {a['attempted_rho_mins']} "
          f"iters {a['solve_iters']} continued {b['selected_continued_probe']}")
    for k in ("trigger", "attempted_rho_mins", "recommended_rho_min", "solve_converged",
              "solver_reported_converged", "solve_iters", "solve_final_rel_residual",
              "compliance", "probe_r50", "probe_r100", "post_solve_fallback_triggered"):
        assert a[k] == b[k], k
    assert a["solve_history"] == b["solve_history"]
    assert torch.equal(a["x"], b["x"])
    kept_first = a["attempted_rho_mins"].split(";")[0] == f"{1e-12:.0e}"
    assert b["selected_continued_probe"] == int(kept_first and a["trigger"] == "keep"
                                                and not a["post_solve_fallback_triggered"])


def test_wait_flags_reported_and_cleared():
    import torch
    from gpu_fem import _lib

    dims = (24, 12, 12)
    r = _runner(dims)
    E = torch.as_tensor(r.moduli(1e-12, _mask(43, dims, 0.35)), device="cuda")
    r.gmg.setup(E)
    ref = r.gmg.apply(r.F)
    ref_v = r.gmg.vcycle(r.F)
    torch.cuda.synchronize()
    for which, name in ((0, "Gram-Schmidt sweep"), (1, "persistent coarsest PCG"),
                        (2, "one-launch line search")):
        r.gmg._set_wait_flag(which)
        with pytest.raises(_lib.GfError, match=name):
            r.gmg.apply(r.F)
        # cleared: the next application runs and gives the same bits
        assert torch.equal(r.gmg.apply(r.F), ref)
        r.gmg._set_wait_flag(which)
        with pytest.raises(_lib.GfError, match=name):
            r.gmg.vcycle(r.F)
        assert torch.equal(r.gmg.vcycle(r.F), ref_v)
    r.gmg._set_wait_flag(0)
    r.gmg._set_wait_flag(1)
    with pytest.raises(_lib.GfError, match="Gram-Schmidt sweep and the persistent coarsest PCG"):
        r.gmg.solve(r.F, tol=1e-6, maxiter=5, restart=300)
    h1: list = []
    x1, it1, _ = r.gmg.solve(r.F, tol=1e-6, maxiter=30, restart=300, history=h1)
    r.gmg.setup(E)
    h2: list = []
    x2, it2, _ = r.gmg.solve(r.F, tol=1e-6, maxiter=30, restart=300, history=h2)
    assert it1 == it2 and h1 == h2 and torch.equal(x1, x2)
    with pytest.raises(ValueError):
        r.gmg._set_wait_flag(3)
