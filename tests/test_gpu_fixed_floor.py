"""Compliance and iteration counts against the reference's fixed-floor
controls, and compliance to 1e-10 where the preconditioner is linear.

The reference's compliance table (run_gmg_fixed_floor_controls.py:82-130,
fixed_floor_control_comparison.csv, RTX 4090, canonical FP64 stack, penal
4.5, FGMRES restart = maxiter = 300, tol 1e-6) runs ONE component set per
preset and every case of that preset in manifest order
(fixed_floor_control_manifest.csv): set-up at each floor, FGMRES from x = 0,
recomputed true residual, compliance F.x.  The fine lambda_max is estimated
at the preset's first set-up and cached for the rest of its session
(multigrid_v4.py:1170-1192), so the cases are replayed in the same order:

  cantilever_gpu_medium (80x40x20): cant_s23_p035, cant_s31_p035, C64_MF
  bridge_gpu_medium (90x30x15):     bridge_s23_p035 (1e-2 only)
  cantilever_gpu_xlarge (160x80x40): C512_MF
  bracket_gpu_500k (80x160x40):     Brk500_MF

Random cases use density = (default_rng(seed).random(n) < p) as 0/1 (void
E = floor exactly); density cases the shipped optimized field clipped to
[0, 1]; E = floor + (1 - floor) rho^4.5.

Bar: the convergence flag equal to the table and every true residual
<= 1e-6; solve_iters equal to the table, except that a count may differ by
one where the crossing of tol is decided by rounding (our residual at the
earlier of the two counts within a factor 3 of tol); compliance within 1e-6
relative of the table (2e-8 and exact iterations for the optimized
densities and the bridge state).  The Bernoulli 0/1 cantilever states of this session
are where the canonical preconditioner's nonlinearity (line-search clipping
and accept/reject, inner FGMRES with tol 0) makes the trajectory depend on
rounding: an independent restatement, the CPU oracle, solves cant_s23_p035
at 1e-12 in 80 iterations (residual 9.79e-7, compliance 10.568883029) where
the table has 81 (10.5688866874), 3.5e-7 apart, and the GPU takes 80
(10.5688816787).  The optimized-density cases and the bridge state are not
rounding-sensitive and match the table to 1e-13..1e-15 in compliance
(C512_MF at 1e-12: 1.3e-8) with equal iterations.  The linear-stack test
below holds compliance to 1e-10, where the mathematics allows it.
"""

import csv
import hashlib
import io
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SESSIONS = [
    ("cantilever_gpu_medium", (80, 40, 20), ["cant_s23_p035", "cant_s31_p035", "C64_MF"]),
    ("bridge_gpu_medium", (90, 30, 15), ["bridge_s23_p035"]),
    ("cantilever_gpu_xlarge", (160, 80, 40), ["C512_MF"]),
    ("bracket_gpu_500k", (80, 160, 40), ["Brk500_MF"]),
]
# fixed_floor_control_manifest.csv: random cases (seed, p); every case uses
# floors 1e-12;1e-3;1e-2 except the bridge (1e-2)
RANDOM = {"cant_s23_p035": (23, 0.35), "cant_s31_p035": (31, 0.35),
          "bridge_s23_p035": (23, 0.35)}
FLOORS = {"bridge_s23_p035": (1e-2,)}
COMPLIANCE_RTOL = 1e-6
# cases that are not rounding-sensitive: exact iterations, compliance 2e-8
STRICT = {"C64_MF", "C512_MF", "Brk500_MF", "bridge_s23_p035"}
STRICT_RTOL = 2e-8


def _table(golden_dir):
    rows = {}
    with open(os.path.join(golden_dir, "fixed_floor_controls.csv")) as f:
        for r in csv.DictReader(f):
            rows[(r["case_id"], float(r["rho_min"]))] = r
    return rows


def _density(golden_dir, case, n):
    if case in RANDOM:
        seed, p = RANDOM[case]
        return (np.random.default_rng(seed).random(n) < p).astype(np.float64)
    if case == "C64_MF":
        rho = np.load(os.path.join(golden_dir, "C64_MF__rho_final.npy"), allow_pickle=False)
    else:
        z = np.load(os.path.join(golden_dir, f"{case}__rho_final.npz"), allow_pickle=False)
        rho = z["rho"]
        buf = io.BytesIO()
        np.save(buf, rho)
        # the reference manifest's SHA-256 of the shipped .npy file
        assert hashlib.sha256(buf.getvalue()).hexdigest() == str(z["npy_sha256"]), case
    assert rho.size == n, case
    return np.clip(np.asarray(rho, dtype=np.float64).reshape(-1), 0.0, 1.0)


@pytest.mark.parametrize("preset,dims,cases", SESSIONS, ids=[s[0] for s in SESSIONS])
def test_fixed_floor_controls_match_reference_table(golden_dir, preset, dims, cases):
    import torch
    from gpu_fem import policy as gpol

    table = _table(golden_dir)
    n = int(np.prod(dims))
    runner = gpol.GuardedFloorPolicy(preset, penal=4.5)
    worst = 0.0
    bad = []
    n_shift = 0
    for case in cases:
        rho = _density(golden_dir, case, n)
        for floor in FLOORS.get(case, (1e-12, 1e-3, 1e-2)):
            ref = table[(case, floor)]
            h: list = []
            _, x, it, rep, ok, fin, _ = runner.solve_at_floor(floor, maxiter=300, history=h,
                                                               rho_field=rho)
            comp = runner_dot(runner, x)
            c_ref = float(ref["compliance"])
            rel = abs(comp - c_ref) / abs(c_ref)
            worst = max(worst, rel)
            print(f"{preset} {case} rho {floor:g}: iters {it} (table {ref['solve_iters']}), "
                  f"residual {fin:.3e} (table {float(ref['final_rel_residual']):.3e}), "
                  f"compliance {comp:.12g} (table {c_ref:.12g}, rel {rel:.2e}), "
                  f"last residuals {['%.3e' % v for v in h[-3:]]}", flush=True)
            tag = (case, floor)
            it_ref = int(float(ref["solve_iters"]))
            if it != it_ref:
                # a one-step difference where the crossing of tol is marginal:
                # h[k] is the residual after k iterations (h[0] = 1)
                k = min(it, it_ref)
                marginal = abs(it - it_ref) == 1 and h[k] <= 3e-6 and case not in STRICT
                n_shift += 1
                if not marginal:
                    bad.append((tag, "iters", it, it_ref))
            if int(ok) != int(ref["solve_converged"]):
                bad.append((tag, "converged", ok, ref["solve_converged"]))
            if not fin <= 1e-6:
                bad.append((tag, "true residual", fin))
            if not rel <= (STRICT_RTOL if case in STRICT else COMPLIANCE_RTOL):
                bad.append((tag, "compliance", rel))
    torch.cuda.synchronize()
    print(f"{preset}: worst compliance rel diff {worst:.2e}; {n_shift} rounding-decided "
          "one-step iteration differences")
    assert not bad, bad


def runner_dot(runner, x):
    from gpu_fem.solver_v2 import dot

    return dot(runner.F, x)


def test_linear_stack_compliance_1e10():
    """With a linear preconditioner (Chebyshev V-cycle, dense coarsest solve,
    no line search, no inner or root-local Krylov) FGMRES is right-
    preconditioned GMRES, a smooth function of its data: the GPU and the
    oracle take the same iterations and their compliance agrees to 1e-10
    relative (the north star's tolerance), the iterates to 1e-8."""
    import torch
    from gpu_fem import policy as gpol
    from gpu_fem.solver_v2 import dot
    from oracle import fem as ofem, gmg as ogmg, krylov as okry

    stack = dict(n_levels=4, fine_smoother="fp64", coarse_correction_policy="none",
                 root_local_correction_mode="none", inner_krylov_steps=0)
    preset, dims, floor = "cantilever_gpu_medium", (80, 40, 20), 1e-3
    rho = (np.random.default_rng(23).random(int(np.prod(dims))) < 0.35).astype(np.float64)
    runner = gpol.GuardedFloorPolicy(preset, penal=4.5, **stack)
    hg: list = []
    _, xg, itg, _, okg, fing, _ = runner.solve_at_floor(floor, maxiter=300, history=hg,
                                                       rho_field=rho)
    cg = dot(runner.F, xg)
    torch.cuda.synchronize()

    bc, free, edof, F, mf = ofem.build_components(ofem.get_preset(preset))
    og = ogmg.GalerkinMatFreeGMG(mf_op=mf, free=free, nelx=dims[0], nely=dims[1], nelz=dims[2],
                                 KE_UNIT=ofem.KE_UNIT_3D, **stack)
    E = floor + (1.0 - floor) * rho ** 4.5
    og.setup(E)
    ho: list = []
    xo, ito, cvo = okry.fgmres(lambda v: mf.matvec(v, E), F, og.apply, tol=1e-6, maxiter=300,
                               restart=300, history=ho)
    co = float(F @ xo)
    xg_h = xg.cpu().numpy()
    rel_c = abs(cg - co) / abs(co)
    rel_x = np.linalg.norm(xg_h - xo) / np.linalg.norm(xo)
    rel_h = max(abs(a - b) / b for a, b in zip(hg, ho))
    print(f"linear stack 80x40x20 rho 1e-3: iters gpu {itg} oracle {ito}; compliance "
          f"{cg:.15g} vs {co:.15g} (rel {rel_c:.2e}); x rel {rel_x:.2e}; history rel {rel_h:.2e}")
    assert itg == ito and okg and cvo
    assert fing <= 1e-6
    assert rel_c <= 1e-10
    assert rel_x <= 1e-8
    assert rel_h <= 1e-8
