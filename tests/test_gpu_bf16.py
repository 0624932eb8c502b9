"""The bf16 matrix-core fine operator (one LDS-tiled kernel) and the bf16
fine smoother built on it (multigrid_v4.py:1765-1780, 1831-1848).

Bars:
* operator: equal to the oracle's restatement of the fused bf16 matvec
  (cuda_fused_matvec.py:138-246: KE and u rounded to bf16, FP32 products
  and sums) to FP32 accumulation rounding, 2e-6 of the largest entry, on
  both tile shapes (TZ = 16 and TZ = 8 are chosen by grid shape) and on
  grids of several tiles in every direction;
* preconditioner: one base cycle with the bf16 smoother against the
  oracle's with the same smoother within 1e-5 of the largest entry, and at
  least ten times closer to it than to the FP32-smoother cycle -- the bf16
  operator is the one in use; the whole application (inner FGMRES, line
  search and root-local correction around the cycle) within 1e-3: those
  nonlinear wrappers amplify the FP32 rounding differences of the smoothing
  recurrences, and a bf16 operand on a rounding boundary may round the other
  way;
* solves: the guarded policy with the bf16 smoother converges to a
  recomputed true residual <= 1e-6 and takes the decisions the FP32 smoother
  takes.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STACK = dict(n_levels=4, coarse_correction_policy="residual_line_search",
             root_local_correction_mode="node_block_inner_fgmres",
             root_local_correction_inner_steps=20, root_local_correction_inner_restart=20,
             inner_krylov_steps=10, inner_krylov_restart=10)


def _t(a):
    import torch

    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _n(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("dims", [(9, 5, 3), (40, 20, 10), (21, 30, 7), (100, 50, 50)])
def test_bf16_operator_against_oracle(dims):
    import time

    import torch
    from gpu_fem.cuda_fused_matvec import FusedMatvec
    from oracle import fem

    spec = fem.cantilever_spec(*dims)
    bc, free, edof, F, mf = fem.build_components(spec)
    rng = np.random.default_rng(7)
    E = np.where(rng.random(edof.shape[0]) < 0.4, 1.0, 1e-3).astype(np.float32)
    u = rng.standard_normal(len(free)).astype(np.float32)
    fm = FusedMatvec(_t(edof.astype(np.int32)), _t(fem.KE_UNIT_3D), bc.ndof, grid=dims)
    yg = _n(fm.matvec(_t(u), _t(E), _t(free.astype(np.int64)), dtype="bf16"))
    t0 = time.perf_counter()
    yo = mf.matvec_bf16(u, E)
    t_o = time.perf_counter() - t0
    err = np.abs(yg - yo).max() / np.abs(yo).max()
    # the FP32 form, to show the two differ by the operand rounding
    yf = _n(fm.matvec(_t(u), _t(E), _t(free.astype(np.int64)), dtype="fp32"))
    gap = np.abs(yg - yf).max() / np.abs(yf).max()
    torch.cuda.synchronize()
    print(f"bf16 operator {dims}: max err vs oracle {err:.2e} of max |y| (oracle {t_o:.2f} s); "
          f"bf16 vs fp32 form {gap:.2e}")
    assert err <= 2e-6
    assert gap >= 1e-4


def _pair(dims, kind, smoother, rho_min=1e-3, seed=0, p=0.35):
    from gpu_fem import policy as gpol
    from oracle import fem, gmg as ogmg, policy as opol

    if kind == "bridge":
        spec_o = fem._bridge(dims, "b", (0.0, 0.25, 0.5))
        from gpu_fem.presets import _bridge
        spec_g = _bridge(*dims, (0.0, 0.25, 0.5))
    else:
        spec_o = fem.cantilever_spec(*dims, load=kind)
        from gpu_fem.presets import cantilever
        spec_g = cantilever(*dims, load=kind)
    bc, free, edof, F, mf = fem.build_components(spec_o)
    st = dict(STACK, fine_smoother=smoother)
    og = ogmg.GalerkinMatFreeGMG(mf_op=mf, free=free, nelx=dims[0], nely=dims[1], nelz=dims[2],
                                 KE_UNIT=fem.KE_UNIT_3D, **st)
    gg = gpol._build_components(spec_g, **st)[7]
    mask = opol.bernoulli_mask(seed, dims, p).ravel()
    E = opol.simp_modulus(np.where(mask, 1.0, rho_min), rho_min, 4.5)
    og.setup(E)
    gg.setup(_t(E))
    return og, gg, free


@pytest.mark.parametrize("dims,kind", [((16, 8, 8), "tip"), ((12, 6, 6), "bridge")])
def test_bf16_smoother_preconditioner_against_oracle(dims, kind):
    og, gg, free = _pair(dims, kind, "bf16")
    assert gg.level_precisions[:2] == ["bf16", "fp32"]
    rhs = np.random.default_rng(11).standard_normal(len(free))
    zo = og.apply(rhs)
    zg = _n(gg.apply(_t(rhs)))
    vo = og._apply_base_cycle(rhs)
    vg = _n(gg.vcycle(_t(rhs)))
    _, g32, _ = _pair(dims, kind, "fp32")
    v32 = _n(g32.vcycle(_t(rhs)))
    e_apply = np.abs(zg - zo).max() / np.abs(zo).max()
    e_cycle = np.abs(vg - vo).max() / np.abs(vo).max()
    gap = np.abs(vg - v32).max() / np.abs(v32).max()
    print(f"bf16 smoother {dims} {kind}: apply {e_apply:.2e}, cycle {e_cycle:.2e} vs oracle; "
          f"bf16 vs fp32 smoother cycle {gap:.2e}")
    assert e_cycle <= 1e-5
    assert gap >= 10 * e_cycle
    assert e_apply <= 1e-3


def test_bf16_hierarchy_without_fused_operator_refuses_to_smooth():
    """multigrid_v4.py:1774-1779: BF16 smoothing without the fused operator
    raises RuntimeError when the hierarchy is first applied."""
    from gpu_fem.multigrid_v4 import GalerkinMatFreeGMG
    from gpu_fem import policy as gpol
    from gpu_fem.presets import cantilever

    comps = gpol._build_components(cantilever(8, 4, 4), **dict(STACK, fine_smoother="fp32"))
    spec, bc, free, free_gpu, edof_gpu, F_gpu, mf_op = comps[:7]
    from oracle import fem
    g = GalerkinMatFreeGMG(mf_op=mf_op, free=free, free_gpu=free_gpu, nelx=8, nely=4, nelz=4,
                           KE_UNIT=fem.KE_UNIT_3D, fine_smoother="bf16", fused_op=None)
    import torch
    g.setup(torch.ones(128, dtype=torch.float64, device="cuda"))
    with pytest.raises(RuntimeError, match="BF16 fine smoothing requires"):
        g.apply(F_gpu)


def test_bf16_smoother_policy_states():
    """Guarded policy states (80x40x20 cantilever, Bernoulli densities, penal
    4.5) with the bf16 and with the FP32 fine smoother: equal decisions and
    accepted floors, true residuals <= 1e-6."""
    import torch
    from gpu_fem import policy as gpol

    dims = (80, 40, 20)
    rows = {}
    for sm in ("fp32", "bf16"):
        runner = gpol.GuardedFloorPolicy("cantilever_gpu_medium", penal=4.5, fine_smoother=sm)
        rows[sm] = []
        for seed, p in ((23, 0.35), (31, 0.35), (5, 0.5)):
            rho = (np.random.default_rng(seed).random(int(np.prod(dims))) < p).astype(np.float64)
            r = runner.run_state(rho_field=rho)
            torch.cuda.synchronize()
            rows[sm].append(r)
            print(f"{sm} seed {seed} p {p}: trigger {r['trigger']} floor "
                  f"{r['recommended_rho_min']:g} attempted {r['attempted_rho_mins']} iters "
                  f"{r['solve_iters']} residual {r['solve_final_rel_residual']:.3e}", flush=True)
        del runner
    for a, b in zip(rows["fp32"], rows["bf16"]):
        for k in ("trigger", "recommended_rho_min", "attempted_rho_mins", "solve_converged"):
            assert a[k] == b[k], k
        if b["solve_converged"]:
            assert b["solve_final_rel_residual"] <= 1e-6
