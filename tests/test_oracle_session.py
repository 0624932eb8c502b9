"""Session replay of the oracle against the reference's held-out history.

The held-out cantilever suite ran as one session, so every state uses the
lambda_max estimated at the session's first set-up (seed 41, p 0.08,
rho_min 1e-12).  Re-creating that set-up and then taking seed 43, p 0.35 at
rho_min 1e-12 through two canonical FGMRES iterations reproduces the
reference's recorded probe residuals (RTX 4090, CuPy) to ~1e-14: this pins
the whole canonical stack -- set-up, lambda cache, contrast adaptation,
Chebyshev recurrence, line searches, root-local and inner FGMRES, outer MGS
and Givens -- at the granularity of single iterations.  About 40 s on the CPU.
"""

import csv
import os

import numpy as np

from oracle import fem, gmg, krylov, policy

STACK = dict(n_levels=4, fine_smoother="fp64", coarse_correction_policy="residual_line_search",
             root_local_correction_mode="node_block_inner_fgmres",
             root_local_correction_inner_steps=20, root_local_correction_inner_restart=20,
             inner_krylov_steps=10, inner_krylov_restart=10)


def test_session_replay_matches_reference_history(golden_dir):
    ref = {}
    with open(os.path.join(golden_dir, "heldout_cantilever_history_subset.csv")) as f:
        for r in csv.DictReader(f):
            if (r["seed"], r["solid_probability"], r["phase"]) == ("43", "0.35", "probe"):
                ref[int(r["iter"])] = float(r["rel_residual"])
    spec = fem.get_preset("cantilever_gpu_medium")
    bc, free, edof, F, mf = fem.build_components(spec)
    g = gmg.GalerkinMatFreeGMG(mf_op=mf, free=free, nelx=80, nely=40, nelz=20,
                               KE_UNIT=fem.KE_UNIT_3D, **STACK)

    def E_of(seed, p):
        m = policy.bernoulli_mask(seed, (80, 40, 20), p).ravel()
        return policy.simp_modulus(np.where(m, 1.0, 1e-12), 1e-12, 4.5)

    g.setup(E_of(41, 0.08))
    lam_first = g.lambda_max_estimate
    E = E_of(43, 0.35)
    g.setup(E)
    assert g.lambda_max_estimate == lam_first          # cached across the session
    h = []
    krylov.fgmres(lambda v: mf.matvec(v, E), F, g.apply, tol=0.0, maxiter=2, restart=300,
                  history=h)
    assert h[0] == 1.0
    for k in (1, 2):
        assert abs(h[k] - ref[k]) <= 1e-12 * ref[k], (k, h[k], ref[k])
