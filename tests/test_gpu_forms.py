"""The added forms at their own workload sizes, through the C ABI.

Batched form (cantilever grids between 30x15x15 and 80x40x20, three floors,
FGMRES restart 100, SIMP p = 3, the unit edge load): concurrent solves over
mixed grid sizes, FP64 and FP32 multigrid.  Two odd-dimension 30x15x15 states
are also solved by the oracle at every floor (restart 100, maxiter 300,
canonical stack): equal FGMRES iteration counts and guard decisions, and
compliance within 1e-6 relative where the guard accepts (the canonical
preconditioner is nonlinear; see test_gpu_fixed_floor for the method's own
spread).  Every accepted solve, in either precision, has a recomputed true
residual <= 1e-6.

Precision form at 1,000,000 elements (200x100x50 filtered-sigmoid density):
a short guarded policy (probe 20, r50 at 10, ladder {1e-2}) with FP64, with
FP32 multigrid and with the bf16 fine smoother (FP32 coarse levels): equal
decisions, and each reduced form's compliance and solution within stated
bounds of the FP64 form's (1e-6 and 1e-4 relative), with its true residual
recomputed in FP64 <= 1e-6.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STACK = dict(n_levels=4, fine_smoother="fp64", coarse_correction_policy="residual_line_search",
             root_local_correction_mode="node_block_inner_fgmres",
             root_local_correction_inner_steps=20, root_local_correction_inner_restart=20,
             inner_krylov_steps=10, inner_krylov_restart=10)


def test_batched_form_range_against_oracle():
    from gpu_fem.batch import BatchedFloorSolver, sample_batch_states
    from oracle import fem as ofem, gmg as ogmg, krylov as okry

    odd = [{"dims": (30, 15, 15), "seed": 1002, "probability": 0.5},
           {"dims": (30, 15, 15), "seed": 1001, "probability": 0.35}]
    states = odd + sample_batch_states(64, seed=0)[:4]
    assert any(d % 2 for s in states for d in s["dims"])
    rows = {}
    for prec in ("fp64", "fp32"):
        rows[prec] = BatchedFloorSolver(workers=4, restart=100, maxiter=300,
                                        precision=prec).run(states)
        assert len(rows[prec]) == 3 * len(states)
        for r in rows[prec]:
            print(f"{prec} {r['dims']} seed {r['seed']} rho {r['rho_min']:g}: iters {r['iters']} "
                  f"accepted {r['accepted']} residual {r['true_rel_residual']:.2e} "
                  f"compliance {r['compliance']:.10g}", flush=True)
            if r["accepted"]:
                assert r["true_rel_residual"] <= 1e-6
    agree = sum(a["accepted"] == b["accepted"] for a, b in zip(rows["fp64"], rows["fp32"]))
    both = [(a, b) for a, b in zip(rows["fp64"], rows["fp32"]) if a["accepted"] and b["accepted"]]
    worst = max(abs(a["compliance"] - b["compliance"]) / abs(a["compliance"]) for a, b in both)
    print(f"fp32 vs fp64: {agree}/{len(rows['fp64'])} guard decisions agree; compliance of "
          f"solves both accept within {worst:.2e}")
    assert worst <= 1e-6

    dims = (30, 15, 15)
    bc, free, edof, F, mf = ofem.build_components(ofem.cantilever_spec(*dims, load="edge"))
    b_norm = np.linalg.norm(F)
    for i, st in enumerate(odd):
        og = ogmg.GalerkinMatFreeGMG(mf_op=mf, free=free, nelx=dims[0], nely=dims[1],
                                     nelz=dims[2], KE_UNIT=ofem.KE_UNIT_3D, **STACK)
        solid = np.random.default_rng(st["seed"]).random(int(np.prod(dims))) < st["probability"]
        for k, rho_min in enumerate((1e-12, 1e-3, 1e-2)):
            E = rho_min + (1.0 - rho_min) * np.where(solid, 1.0, rho_min) ** 3.0
            og.setup(E)
            x, it, _ = okry.fgmres(lambda v: mf.matvec(v, E), F, og.apply, tol=1e-6,
                                   maxiter=300, restart=100)
            res = np.linalg.norm(F - mf.matvec(x, E)) / b_norm
            g = rows["fp64"][3 * i + k]
            assert g["rho_min"] == rho_min and g["seed"] == st["seed"]
            c = float(F @ x)
            rel = abs(g["compliance"] - c) / abs(c)
            print(f"oracle seed {st['seed']} rho {rho_min:g}: iters {it} (gpu {g['iters']}), "
                  f"accepted {int(res <= 1e-6)} (gpu {g['accepted']}), residual {res:.2e}, "
                  f"compliance rel {rel:.2e}", flush=True)
            assert g["iters"] == it
            assert g["accepted"] == int(res <= 1e-6)
            if g["accepted"]:
                assert rel <= 1e-6


def test_precision_form_1m():
    import torch
    from gpu_fem import policy as gpol
    from gpu_fem.batch import precision_stack
    from gpu_fem.density_field import filtered_sigmoid_density
    from gpu_fem.presets import cantilever

    dims = (200, 100, 50)
    rho = filtered_sigmoid_density(0, dims, radius=3.0, sharpness=8.0, volfrac=0.3)
    spec = cantilever(*dims, load="edge")
    kw = dict(penal=3.0, baseline_rho_min=1e-12, raised_rho_mins=(1e-2,), probe_iters=20,
              probe_r50=10, restart=300, maxiter=300, tol=1e-6)
    rows, xs = {}, {}
    for prec in ("fp64", "fp32", "bf16"):
        runner = gpol.GuardedFloorPolicy(spec, **kw, **precision_stack(prec))
        row = runner.run_state(rho_field=rho, keep_x=True)
        torch.cuda.synchronize()
        xs[prec] = row.pop("x")
        rows[prec] = row
        print(f"1M {prec}: trigger {row['trigger']} floor {row['recommended_rho_min']:g} "
              f"iters {row['solve_iters']} residual {row['solve_final_rel_residual']:.3e} "
              f"compliance {row['compliance']:.12g}", flush=True)
        del runner
        torch.cuda.empty_cache()
    a = rows["fp64"]
    assert a["solve_converged"] == 1
    for prec in ("fp32", "bf16"):
        b = rows[prec]
        for k in ("trigger", "recommended_rho_min", "attempted_rho_mins", "solve_converged"):
            assert a[k] == b[k], (prec, k)
        assert b["solve_final_rel_residual"] <= 1e-6
        c_err = abs(b["compliance"] - a["compliance"]) / abs(a["compliance"])
        x_err = float(torch.linalg.vector_norm(xs[prec] - xs["fp64"])
                      / torch.linalg.vector_norm(xs["fp64"]))
        print(f"1M {prec} vs fp64: compliance rel {c_err:.2e}, solution rel {x_err:.2e}, "
              f"iterations {b['solve_iters']} vs {a['solve_iters']}")
        assert c_err <= 1e-6
        assert x_err <= 1e-4
