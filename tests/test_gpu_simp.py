"""In-loop SIMP solve (SolverV4 matrix-free GMG path, solver_v4.py:149-282)
and its guarded policy wrapper (run_simp_floor_trajectory.py:65-224) on the
GPU, against the oracle."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(dims, seed=5):
    from oracle import fem

    spec = fem.cantilever_spec(*dims)
    bc, free, edof, F, mf = fem.build_components(spec)
    rng = np.random.default_rng(seed)
    return spec, bc, free, edof, F, mf, rng


def _oracle_solve(mf, free, edof, F, dims, rho, penal, Emin, tol, maxiter, x0=None):
    from oracle import fem, gmg as ogmg, krylov as okry

    E = Emin + (1 - Emin) * rho ** penal
    g = ogmg.GalerkinMatFreeGMG(mf_op=mf, free=free, nelx=dims[0], nely=dims[1], nelz=dims[2],
                                KE_UNIT=fem.KE_UNIT_3D, n_levels=4, coarse_smooth_iters=2,
                                fine_smoother="fp64", fine_smoother_degree=2)
    g.setup(E, fine_diag=mf.extract_diagonal(E))
    h = []
    x, it, ok = okry.pcg(lambda v: mf.matvec(v, E), F, g.apply, x0=x0, tol=tol, maxiter=maxiter,
                         history=h)
    full = np.zeros(3 * np.prod([d + 1 for d in dims]))
    full[free] = x
    dc = -penal * (1 - Emin) * rho ** (penal - 1) * fem.element_energy(edof, fem.KE_UNIT_3D, full)
    return float(F @ x), dc, it, x


def test_solverv4_compliance_sensitivity_and_warm_start():
    from gpu_fem.solver_v4 import SolverV4
    from oracle import fem

    dims = (12, 6, 6)
    spec, bc, free, edof, F, mf, rng = _case(dims)
    rho = rng.uniform(0.3, 1.0, edof.shape[0])
    s = SolverV4(edof, None, None, fem.KE_UNIT_3D, free, bc.F, bc.ndof, cg_tol=1e-8,
                 cg_maxiter=200, grid_dims=dims, gmg_fine_smoother="fp64")
    c_g, dc_g = s.solve(rho, 3.0)
    assert s.last_outer_solver == "pcg"
    c_o, dc_o, it_o, x_o = _oracle_solve(mf, free, edof, F, dims, rho, 3.0, 1e-9, 1e-8, 200)
    assert s.last_cg_iters == it_o
    assert abs(c_g - c_o) <= 1e-10 * abs(c_o)
    assert np.abs(dc_g - dc_o).max() <= 1e-9 * np.abs(dc_o).max()
    assert s.last_true_rel_residual <= 1e-8
    # second SIMP iteration: warm start from the previous displacement
    rho2 = np.clip(rho + 0.02 * rng.standard_normal(rho.size), 0.2, 1.0)
    c_g2, dc_g2 = s.solve(rho2, 3.0)
    c_o2, dc_o2, it_o2, _ = _oracle_solve(mf, free, edof, F, dims, rho2, 3.0, 1e-9, 1e-8, 200,
                                          x0=x_o)
    assert s.last_cg_iters == it_o2
    assert abs(c_g2 - c_o2) <= 1e-9 * abs(c_o2)
    assert np.abs(dc_g2 - dc_o2).max() <= 1e-8 * np.abs(dc_o2).max()


def test_guarded_in_loop_policy():
    from gpu_fem.solver_v4 import GuardedAdaptiveFloorSolverV4
    from oracle import fem

    dims = (16, 8, 8)
    spec, bc, free, edof, F, mf, rng = _case(dims)
    kw = dict(cg_tol=1e-6, cg_maxiter=200, grid_dims=dims, gmg_fine_smoother="fp64",
              gmg_outer_solver="fgmres", gmg_restart=100)
    s = GuardedAdaptiveFloorSolverV4(edof, None, None, fem.KE_UNIT_3D, free, bc.F, bc.ndof, **kw)
    # full material: the probe itself converges -> probe_converged_keep
    c1, dc1 = s.solve(np.ones(edof.shape[0]), 3.0)
    ev = s.policy_events[-1]
    assert ev["trigger"] == "probe_converged_keep" and ev["attempt_count"] == 1
    assert ev["accepted_true_rel_residual"] <= 1e-6
    # a 0/1 layout with 30 % void: the policy raises or guards; every accepted
    # solve meets the tolerance and the floor sequence follows the rules
    rho = np.where(rng.random(edof.shape[0]) < 0.7, 1.0, 0.0)
    c2, dc2 = s.solve(rho, 3.0)
    ev = s.policy_events[-1]
    att = ev["attempted_rho_mins"].split(";")
    assert att[0] == "1e-12"
    assert ev["accepted_true_rel_residual"] <= 1e-6
    assert float(att[-1]) == ev["selected_rho_min"]
    if ev["trigger"] in ("high_r50", "plateau"):
        assert att[1:] == ["1e-03", "1e-02"][: len(att) - 1]
    elif ev["trigger"] == "keep":
        assert att == ["1e-12", "1e-12"]
    assert np.isfinite(c2) and np.all(np.isfinite(dc2)) and np.all(dc2 <= 0)
    # mostly void: no floor of the ladder converges within the budget -> the
    # policy records a failed event and raises, as the reference does
    rho = np.where(rng.random(edof.shape[0]) < 0.1, 1.0, 0.0)
    with pytest.raises(RuntimeError, match="exhausted the ladder"):
        s.solve(rho, 3.0)
    ev = s.policy_events[-1]
    assert ev["fallback_used"] == 1 and np.isnan(ev["selected_rho_min"])
    assert ev["attempted_rho_mins"].split(";")[-1] == "1e-02"


def test_batched_mixed_sizes_match_single_solves():
    """BatchedFloorSolver: concurrent states of different grids give the same
    rows as solving each state alone (deterministic kernels), and every
    accepted solve passes the recomputed-residual guard."""
    from gpu_fem.batch import BatchedFloorSolver, sample_batch_states

    states = sample_batch_states(5, seed=3, lo=(10, 5, 5), hi=(20, 10, 9))
    dims = [s["dims"] for s in states]
    assert len(set(dims)) > 1 and any(d % 2 for s in dims for d in s)
    par = BatchedFloorSolver(workers=3, restart=40, maxiter=120, precision="fp32").run(states)
    seq = BatchedFloorSolver(workers=1, restart=40, maxiter=120, precision="fp32").run(states)
    assert len(par) == len(seq) == 3 * len(states)
    for a, b in zip(par, seq):
        for k in ("dims", "rho_min", "iters", "accepted", "true_rel_residual", "compliance"):
            assert a[k] == b[k], (k, a[k], b[k])
        if a["accepted"]:
            assert a["true_rel_residual"] <= 1e-6
