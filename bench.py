"""Benchmark: accepted guarded solves/s of the plain FP64 GMG-FGMRES policy.

One step = one density state taken through probe -> classify -> floor ladder
-> recomputed-residual guard on one MI355X (the unit of the reference's
headline number, recorded_policy_time_s).  Workload (default): cantilever
80x40x20 (64,000 hex8 elements, 206,640 free DOFs), E0=1, nu=0.3, SIMP p=3,
x=0 face clamped, unit downward load split over the x=L, y=0 edge nodes;
Bernoulli densities p in {0.2, 0.35, 0.5} over seeds 0..31 (taken in
seed-major order: 96 states, the first 12 a warm-up and the other 84
timed by default), ladder {1e-12 probe, 1e-3, 1e-2}, probe 100 / r50 at 50,
thresholds 1e-2 / 1e-4 / 0.6, FGMRES restart=maxiter=300, tol 1e-6.

Prints ONE JSON line (the driver contract) with a roofline object for the
fine matvec (timed live with HIP events on the solver's stream) and a CPU
baseline (the NumPy oracle on a bounded sample, scaled per state).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "solver-admissibility-and-guarded-floor-selection_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# HBM bytes per FP64 matvec launch on the 80x40x20 grid (LDS-tiled kernel),
# from rocprofv3 --pmc FETCH_SIZE (1302.9 KiB) and WRITE_SIZE (1615.0 KiB) per
# dispatch, each scaled by what the counter reads for a 64 MiB device copy
# (FETCH 32783.4 KiB -> x1.99906; WRITE 65536 KiB -> x1.0): 2.67 MB read +
# 1.65 MB written against 2.17 + 1.65 MB algorithmic
# (profiles/fine_matvec_fp64_80x40x20_roofline.md)
TRAFFIC_BYTES = 4320893
# HBM bytes per level-1 coarse-stencil launch (80x40x20 hierarchy, rotated
# over hierarchies beyond the Infinity Cache), from rocprofv3 --pmc FETCH_SIZE
# (9287.06 KiB) and WRITE_SIZE (216.78 KiB) per dispatch, scaled by the 64 MiB
# copy's readings (x1.99929, x1.0): 1.05x the 18,406,080 algorithmic bytes
# (profiles/coarse_stencil_80x40x20_roofline.md)
STENCIL_TRAFFIC_BYTES = 19235215
SHAPE_KERNELS = {1: "k_fine_matvec_node", 4: "k_fine_matvec_wave4",
                 8: "k_fine_matvec_tile<TZ=8>", 16: "k_fine_matvec_tile<TZ=16>",
                 72: "k_fine_matvec_mfma<TZ=8>", 80: "k_fine_matvec_mfma<TZ=16>",
                 48: "k_fine_matvec_box<4x8x8 nodes, 1 lane per node>",
                 50: "k_fine_matvec_box<2x8x16 nodes, 1 lane per node>",
                 44: "k_fine_matvec_box<2x8x8 nodes, 2 lanes per node>"}


def progress(msg):
    """One progress line on stderr (the JSON result line stays alone on stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=84,
                    help="timed states (default: every state after the warm-up)")
    ap.add_argument("--warmup", type=int, default=12)
    ap.add_argument("--grid", default="80x40x20")
    ap.add_argument("--penal", type=float, default=3.0)
    ap.add_argument("--load", default="edge", choices=["edge", "tip"])
    ap.add_argument("--probabilities", default="0.2,0.35,0.5")
    ap.add_argument("--seeds", default=",".join(str(s) for s in range(32)))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="policy line only: skip the one-instance, learned and reference-"
                         "parametrisation runs and the roofline loops (for profiling runs)")
    ap.add_argument("--workers", type=int, default=12,
                    help="solver instances running states concurrently (one session)")
    ap.add_argument("--mgs-sweep", type=int, default=1, choices=[0, 1, 2],
                    help="outer FGMRES orthogonalisation: 1 persistent sweep, 2 cooperative "
                         "launch, 0 one launch per projection")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="HIP hardware queues for this process (GPU_MAX_HW_QUEUES, <= 32; "
                         "0 keeps the runtime default; --form precision: 1)")
    ap.add_argument("--states-csv", default="")
    ap.add_argument("--form", default="policy",
                    choices=["policy", "batched", "precision", "learned", "coarsest"],
                    help="policy: the plain FP64 guarded policy on the 80x40x20 Bernoulli "
                         "states (default line); batched: fixed-floor guarded solves of a "
                         "mixed-size batch, FP32 vs FP64 preconditioner; precision: one "
                         "filtered-sigmoid state through the policy with an FP32 vs FP64 "
                         "multigrid (use --grid 200x100x50); learned: train the floor "
                         "network and compare learned + FP32 multigrid with the plain policy; "
                         "coarsest: the coarsest-level PCG in two-launch vs persistent form "
                         "through preconditioner applies (use --grid 200x100x50)")
    ap.add_argument("--density-seed", type=int, default=0)
    ap.add_argument("--floors", default="1e-12,1e-2",
                    help="precision form: probe floor, then the raised ladder")
    ap.add_argument("--probe-iters", type=int, default=100)
    ap.add_argument("--restart", type=int, default=300)
    ap.add_argument("--maxiter", type=int, default=300)
    ap.add_argument("--train-states", type=int, default=768)
    ap.add_argument("--train-grids", default="40x20x20,64x32x16,80x40x20")
    ap.add_argument("--train-seed", type=int, default=1)
    ap.add_argument("--eval-states", type=int, default=96)
    ap.add_argument("--eval-seed", type=int, default=2)
    ap.add_argument("--eval-kinds", default="bernoulli,sigmoid")
    ap.add_argument("--eval-volfrac", type=float, default=0.0,
                    help="fix the filtered-sigmoid volume fraction of the evaluation states")
    ap.add_argument("--eval-workers", type=int, default=0,
                    help="concurrent instances of the evaluation (0: --workers)")
    ap.add_argument("--learned-variants", default="fp64,fp32,fp32_gated",
                    help="comma list of fp64, fp32, fp32_gated, fp64_gated")
    ap.add_argument("--eval-warmup", type=int, default=1, choices=[0, 1])
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--labels", default="",
                    help="label CSV to train from (default: the shipped gpu_fem/data labels)")
    ap.add_argument("--relabel", action="store_true",
                    help="label fresh training states with the plain policy instead")
    ap.add_argument("--confidence", default="auto",
                    help="confidence gate of the gated learned form: a probability, or auto "
                         "(the smallest one whose training-set predictions are 99%% correct)")
    ap.add_argument("--labels-out", default="", help="write the training labels here")
    ap.add_argument("--model", default="",
                    help="learned form: evaluate this trained network (.npz with its training "
                         "grids and per-grid confidence gates) instead of training one")
    ap.add_argument("--theta-out", default="",
                    help="save the trained parameters (.npz: with the training grids)")
    ap.add_argument("--batch-states", type=int, default=64)
    ap.add_argument("--roofline-only", action="store_true",
                    help="time only the fine matvec loop (for rocprofv3 runs)")
    ap.add_argument("--coarse-form", type=int, default=1, choices=[0, 1, 16, 32, 64],
                    help="coarse-level operator form (gf_set_coarse_form): 1 node-per-lane "
                         "workgroups sized by level (default), 16/32/64 that form with the "
                         "node count fixed, 0 one lane per (node, slot)")
    ap.add_argument("--ls-form", type=int, default=1, choices=[0, 1],
                    help="residual line search of the coarse corrections (gf_set_ls_form): "
                         "1 one persistent launch where it fits (default), 0 three launches")
    ap.add_argument("--fine-shapes", default="",
                    help="with --roofline-only (fp64): comma-separated launch shapes to time "
                         "one after the other (1, 4, 8, 16, 72, 80, 48, 50, 44; default: the autotuned one)")
    ap.add_argument("--roofline-kernel", default="fp64", choices=["fp64", "bf16", "stencil"],
                    help="with --roofline-only: the FP64 fine operator, its BF16 MFMA form, or "
                         "the level-1 coarse stencil rotated over set-up hierarchies")
    return ap.parse_args()


def states(args):
    seeds = [int(s) for s in args.seeds.split(",") if s.strip()]
    probs = [float(p) for p in args.probabilities.split(",") if p.strip()]
    while True:
        for s in seeds:
            for p in probs:
                yield s, p


def matvec_roofline(runner, n_launch=200, min_bytes=300 << 20, shape=0):
    """Average duration of the FP64 fine matvec kernel on the solver's stream
    (HIP events), and its algorithmic bytes / flops per launch.  The launches
    rotate over enough (E, u, y) buffer sets that together they exceed the
    256 MB Infinity Cache (at least 5 sets), so each launch reads its inputs
    from HBM.  ``shape`` selects a launch shape (0: the autotuned one)."""
    import torch

    from gpu_fem import _lib

    mf = runner.mf_op
    n_free = mf._n_free
    n_elem = mf._n_elem
    bytes_alg = 8 * n_elem + 2 * 8 * n_free
    n_sets = max(5, -(-min_bytes // bytes_alg))
    sets = []
    for _ in range(n_sets):
        E = torch.rand(n_elem, dtype=torch.float64, device="cuda") + 1e-3
        u = torch.randn(n_free, dtype=torch.float64, device="cuda")
        sets.append((E, u, torch.empty_like(u)))
    st = torch.cuda.current_stream()
    lib = _lib.load()
    prev = mf.launch_shape
    if shape:
        mf.set_launch_shape(shape)
    calls = [(mf.ctx_handle, _lib.ptr(E), _lib.ptr(u), _lib.ptr(y), st.cuda_stream)
             for E, u, y in sets]
    for k in range(2 * n_sets):
        lib.gf_matvec_f64(*calls[k % n_sets])
    st.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(st)
    for k in range(n_launch):
        lib.gf_matvec_f64(*calls[k % n_sets])
    t1.record(st)
    st.synchronize()
    if shape:
        mf.set_launch_shape(prev)
    avg_us = t0.elapsed_time(t1) * 1e3 / n_launch
    flops = 1152 * n_elem
    gbs = bytes_alg / (avg_us * 1e-6) / 1e9
    tfs = flops / (avg_us * 1e-6) / 1e12
    return avg_us, bytes_alg, flops, gbs, tfs


def stencil_roofline(session, spec, stack, n_launch=600, min_bytes=300 << 20):
    """Average duration of the level-1 coarse operator kernel (the 27-point
    3x3-block stencil, plain epilogue: the k_stencil_nodes family holds the
    largest share of the policy's kernel time), launched round-robin over
    enough set-up hierarchies that the stencils they read exceed the 256 MB
    Infinity Cache (so every launch streams its coefficients from HBM), with
    HIP events on the launching stream.  Algorithmic bytes per launch: the
    27 x 9 coefficients of every free node row (1944 B), its three input
    values (24 B, each read once) and three outputs (24 B)."""
    import numpy as np
    import torch

    from gpu_fem import _lib
    from gpu_fem import policy as gpol

    gmgs = [r.gmg for r in session.runners]
    n1 = int(gmgs[0]._n_free[1])
    rows = n1 // 3
    per = rows * 1944
    extra = []
    rng = np.random.default_rng(5)
    n_e = session.runners[0].n_elem
    while per * (len(gmgs) + len(extra)) < min_bytes:
        comps = gpol._build_components(spec, **stack)
        g = comps[7]
        g.setup(torch.as_tensor(np.where(rng.random(n_e) < 0.35, 1.0, 1e-3), device="cuda"))
        extra.append(comps)
    hs = gmgs + [c[7] for c in extra]
    xs = [torch.randn(n1, dtype=torch.float64, device="cuda") for _ in hs]
    ys = [torch.empty_like(x) for x in xs]
    lib = _lib.load()
    st = torch.cuda.current_stream()
    calls = [(h._h, 1, _lib.ptr(x), _lib.ptr(y), st.cuda_stream) for h, x, y in zip(hs, xs, ys)]
    for k in range(2 * len(calls)):
        lib.gf_gmg_level_matvec(*calls[k % len(calls)])
    st.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(st)
    for k in range(n_launch):
        lib.gf_gmg_level_matvec(*calls[k % len(calls)])
    t1.record(st)
    st.synchronize()
    avg_us = t0.elapsed_time(t1) * 1e3 / n_launch
    bytes_alg = rows * (1944 + 48)
    del extra
    return {"avg_launch_us": avg_us, "algorithmic_bytes": bytes_alg, "rows": rows,
            "hierarchies_rotated": len(hs), "coefficient_bytes_rotated": per * len(hs),
            "GB_s": bytes_alg / (avg_us * 1e-6) / 1e9}


HELDOUT_SEEDS = (41, 43, 47, 53, 59, 61, 67, 71)
HELDOUT_P = (0.08, 0.1, 0.12, 0.15, 0.18, 0.2, 0.25, 0.3, 0.35)
REFERENCE_RATE = 0.0282  # accepted guarded solves/s, RTX 4090, the same 72 states


def reference_parametrisation(args):
    """The reference's own held-out cantilever suite (80x40x20 preset, SIMP
    p = 4.5, tip point load, seeds 41..71 x p 0.08..0.35: 72 states) as one
    session on --workers concurrent instances, timed after the session's
    first state; the reference took 2,550.6 s on an RTX 4090 (0.0282
    accepted solves/s)."""
    import numpy as np
    import torch

    from gpu_fem import policy as gpol

    n = 80 * 40 * 20
    sts = [{"solid": np.random.default_rng(s).random(n) < p, "seed": s, "probability": p}
           for s in HELDOUT_SEEDS for p in HELDOUT_P]
    progress(f"reference parametrisation: {len(sts)} held-out states")
    session = gpol.PolicySession("cantilever_gpu_medium", workers=args.workers, penal=4.5,
                                 mgs_sweep=args.mgs_sweep)
    session.run(sts[:1])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    got = session.run(sts[1:])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    acc = sum(int(r["solve_converged"]) for r in got)
    del session
    torch.cuda.empty_cache()
    return {"workload": ("reference held-out suite: cantilever_gpu_medium 80x40x20, SIMP p=4.5, "
                         "tip load, seeds 41-71 x p 0.08-0.35, one session (lambda from seed 41 "
                         "p 0.08, untimed), states 2-72 timed"),
            "states": len(got), "accepted": acc, "wall_s": wall, "value": acc / wall,
            "reference_rtx4090_value": REFERENCE_RATE,
            "ratio_vs_reference": (acc / wall) / REFERENCE_RATE,
            "max_true_rel_residual": max(r["solve_final_rel_residual"] for r in got
                                         if r["solve_converged"])}


def learned_on_states(args, spec, kw, dims, warm, timed, plain_rows, wall_plain, wall_seq):
    """The learned policy with the shipped network (gpu_fem/data) on the
    timed states, --workers concurrent instances: with FP32 multigrid
    ungated (every in-range state solved at the proposed floor) and gated
    (states below the network's calibrated confidence go through the plain
    probe), and gated with the plain policy's FP64 stack (a probed state is
    then solved exactly as the plain line solves it).  Each runs as its own
    session after a warm-up on the same warm-up states as the plain line;
    the pooling and the network's forward pass are inside the timed
    region."""
    import numpy as np
    import torch

    from gpu_fem import learned as L
    from gpu_fem import policy as gpol
    from gpu_fem.batch import precision_stack

    model = L.FloorCNN.load()
    out = {"model": os.path.relpath(L.DEFAULT_MODEL, ROOT),
           "trained_grids": [list(g) for g in (model.trained_grids or [])],
           "in_training_range": model.in_range(dims)}
    acc_plain = sum(int(r["solve_converged"]) for r in plain_rows)

    def predict(sts):
        X = torch.stack([L.pool_density(st["solid"].astype(np.float64), dims) for st in sts])
        feat = torch.as_tensor(np.stack([L.grid_features(dims)] * len(sts)), device="cuda")
        return model.predict_proba(X, feat)

    tau = model.gate_for(dims)
    for name, gate, prec in (("fp32_ungated", 0.0, "fp32"), ("fp32_gated", tau, "fp32"),
                             ("fp64_gated", tau, "fp64")):
        progress(f"learned ({name}, gate {gate:.3f}): warm-up and {len(timed)} timed states")
        session = gpol.PolicySession(spec, workers=args.workers, runner_cls=L.LearnedFloorPolicy,
                                     confidence_threshold=gate, **kw, **precision_stack(prec))

        def attach(sts):
            cls, r50, conf = predict(sts)
            return [dict(st, extra={"predicted_class": int(c), "predicted_r50": float(r),
                                    "predicted_confidence": float(q),
                                    "out_of_range": not model.in_range(dims)})
                    for st, c, r, q in zip(sts, cls, r50, conf)]

        if warm:
            session.run(attach(warm))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        got = session.run(attach(timed))
        torch.cuda.synchronize()
        w = time.perf_counter() - t0
        c = L.compare(plain_rows, got)
        acc = c["learned_accepted"]
        c.update(value=acc / w, wall_s=w, confidence_gate=gate,
                 speedup_vs_plain_fp64_one_instance=(acc / w) / (acc_plain / wall_seq),
                 speedup_vs_plain_fp64_concurrent=(acc / w) / (acc_plain / wall_plain),
                 mean_total_iters=float(np.mean([r["total_iters"] for r in got])))
        out[name] = c
        del session
        torch.cuda.empty_cache()
    return out


def fused_matvec_rates(runner, n_launch=200):
    """Average launch duration (HIP events on the launching stream) of the
    fused matvec's FP32 form (the FP32 owner-computes operator) and of its
    BF16 matrix-core form (one LDS-tiled kernel, k_fine_matvec_bf16) on the
    full DOF layout of the runner's grid.  Algorithmic HBM bytes per launch:
    the FP32 moduli read once (4 B per element), u read and y written once
    (2 x 4 B per DOF); useful flops 1152 per element (24 x 24 multiply-adds)."""
    import torch

    from gpu_fem.cuda_fused_matvec import FusedMatvec

    mf = runner.mf_op
    fm = FusedMatvec(mf._edof_gpu, mf._KE_unit, mf._ndof, grid=mf.grid)
    E = torch.rand(mf._n_elem, dtype=torch.float32, device="cuda") + 1e-3
    u = torch.randn(mf._ndof, dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream()
    u_full, y_full = fm._scratch(u.device)
    u_full.copy_(u)
    bytes_alg = 4 * mf._n_elem + 2 * 4 * mf._ndof
    flops = 1152 * mf._n_elem
    out = {"algorithmic_bytes": bytes_alg, "flops": flops}
    for dt in ("fp32", "bf16"):
        for _ in range(5):
            fm._launch(dt, u_full, y_full, E)
        st.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(st)
        for _ in range(n_launch):
            fm._launch(dt, u_full, y_full, E)
        t1.record(st)
        st.synchronize()
        us = t0.elapsed_time(t1) * 1e3 / n_launch
        out[dt] = {"avg_us": us, "hbm_GB_s": bytes_alg / (us * 1e-6) / 1e9,
                   "tflops": flops / (us * 1e-6) / 1e12}
    out["bf16"]["kernel"] = "k_fine_matvec_bf16 (gf_matvec_full_bf16)"
    out["fp32"]["kernel"] = "fine operator, FP32 (gf_matvec_full_f32)"
    return out


def cpu_baseline(args, dims, gpu_rows):
    """The NumPy oracle on the box's host cores: one set-up and two canonical
    FGMRES iterations on the first benchmark state at the probe floor
    (about 20-40 s), scaled by the GPU run's mean iterations and set-ups per
    accepted state.  Returns the JSON object."""
    import numpy as np

    from oracle import fem, gmg as ogmg, krylov as okry, policy as opol

    spec = fem.cantilever_spec(*dims, load=args.load)
    bc, free, edof, F, mf = fem.build_components(spec)
    stack = dict(n_levels=4, fine_smoother="fp64", coarse_correction_policy="residual_line_search",
                 root_local_correction_mode="node_block_inner_fgmres",
                 root_local_correction_inner_steps=20, root_local_correction_inner_restart=20,
                 inner_krylov_steps=10, inner_krylov_restart=10)
    g = ogmg.GalerkinMatFreeGMG(mf_op=mf, free=free, nelx=dims[0], nely=dims[1], nelz=dims[2],
                                KE_UNIT=fem.KE_UNIT_3D, **stack)
    seed, p = gpu_rows[0]["seed"], gpu_rows[0]["solid_probability"]
    mask = opol.bernoulli_mask(seed, dims, p).ravel()
    E = opol.simp_modulus(np.where(mask, 1.0, 1e-12), 1e-12, args.penal)
    t0 = time.perf_counter()
    g.setup(E)
    t_setup = time.perf_counter() - t0
    # per-iteration cost of the canonical outer FGMRES on this state: time
    # two outer iterations (the sample), scale to the GPU run's per-state work
    t1 = time.perf_counter()
    okry.fgmres(lambda x: mf.matvec(x, E), F, g.apply, tol=0.0, maxiter=2, restart=300)
    per_iter = (time.perf_counter() - t1) / 2.0
    acc = [r for r in gpu_rows if r["solve_converged"]]
    it_per_state = float(np.mean([r["total_iters"] for r in acc])) if acc else float("nan")
    setups_per_state = float(np.mean([r["n_setups"] for r in acc])) if acc else float("nan")
    sec_per_state = it_per_state * per_iter + setups_per_state * t_setup
    # one complete guarded state of the CPU-runnable configuration (cantilever
    # 40x20x20, same parametrisation), timed end to end on the oracle: probe,
    # classification, ladder, guard, in a fresh session (lambda estimated at
    # its first set-up), as the reference runs a state
    c1 = (40, 20, 20)
    spec1 = fem.cantilever_spec(*c1, load=args.load)
    bc1, free1, _, F1, mf1 = fem.build_components(spec1)
    g1 = ogmg.GalerkinMatFreeGMG(mf_op=mf1, free=free1, nelx=c1[0], nely=c1[1], nelz=c1[2],
                                 KE_UNIT=fem.KE_UNIT_3D, **stack)

    def solve1(E1, mx, hist):
        g1.setup(E1)
        return okry.fgmres(lambda x: mf1.matvec(x, E1), F1, g1.apply, tol=1e-6, maxiter=mx,
                           restart=300, history=hist)

    mask1 = opol.bernoulli_mask(0, c1, 0.35).ravel()
    t2 = time.perf_counter()
    row1 = opol.run_policy(mask1, solve1, F1, lambda E1: (lambda x: mf1.matvec(x, E1)),
                           penal=args.penal)
    c1_s = time.perf_counter() - t2
    try:
        import threadpoolctl

        info = threadpoolctl.threadpool_info()
        cores = max([i.get("num_threads", 1) for i in info] + [1])
    except Exception:
        cores = os.cpu_count() or 1
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")),
                             "")
    except OSError:
        pass
    return {
        "value": 1.0 / sec_per_state if sec_per_state > 0 else None,
        "unit": "accepted guarded solves/s",
        "cores": int(cores),
        "kind": "port",
        "sample": (f"oracle (NumPy/SciPy) on state seed={seed} p={p} {dims[0]}x{dims[1]}x{dims[2]}: "
                   f"set-up {t_setup:.2f} s, one canonical FGMRES iteration {per_iter:.2f} s "
                   f"(measured), scaled by the GPU run's {it_per_state:.1f} iterations and "
                   f"{setups_per_state:.2f} set-ups per accepted state (the reference's count: "
                   "probe, every ladder attempt and the selected solve from scratch)"),
        "cpu_model": cpu_model,
        "complete_state_40x20x20": {
            "state": "seed 0, p 0.35, cantilever 40x20x20 (16,000 hex8), same parametrisation",
            "seconds": c1_s, "accepted": int(row1["solve_converged"]),
            "solves_per_s": int(row1["solve_converged"]) / c1_s,
            "trigger": row1["trigger"], "recommended_rho_min": row1["recommended_rho_min"],
            "solve_iters": row1["solve_iters"],
        },
    }


def batched_form(args):
    """Batched form: 64 seeded states on cantilever grids drawn per axis from
    30x15x15 .. 80x40x20 (uneven sizes, odd dims), Bernoulli p cycling over
    {0.2, 0.35, 0.5}, each solved at floors {1e-12, 1e-3, 1e-2} (192 solves)
    with FGMRES(100), maxiter 300, tol 1e-6, accepted only through the
    recomputed-residual guard.  The FP32-preconditioner form and the plain
    FP64 form run the same batch; both lines of results are reported."""
    import numpy as np
    import torch

    from gpu_fem.batch import BatchedFloorSolver, sample_batch_states

    states = sample_batch_states(args.batch_states)
    out = {}
    for prec in ("fp32", "fp64"):
        solver = BatchedFloorSolver(workers=args.workers, restart=100, maxiter=300, tol=1e-6,
                                    penal=args.penal, load=args.load, precision=prec)
        progress(f"batched {prec}: warm-up")
        solver.run(states[: min(len(states), args.workers)])      # warm-up (kernels, pools)
        progress(f"batched {prec}: timed batch of {len(states)} states")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rows = solver.run(states)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        out[prec] = (rows, wall)
    r32, w32 = out["fp32"]
    r64, w64 = out["fp64"]
    acc32 = sum(r["accepted"] for r in r32)
    acc64 = sum(r["accepted"] for r in r64)
    both = [(a, b) for a, b in zip(r32, r64) if a["accepted"] and b["accepted"]]
    crel = [abs(a["compliance"] - b["compliance"]) / abs(b["compliance"]) for a, b in both]
    line = {
        "metric": "accepted guarded solves/s, batched form (FP32 preconditioner) vs plain FP64",
        "value": acc32 / w32, "unit": "accepted guarded solves/s", "n_gpus": 1,
        "steps": len(r32), "warmup": min(len(states), args.workers),
        "ms_per_step": w32 / len(r32) * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 outer / f32 preconditioner",
        "data": "synthetic (seeded Bernoulli density states)",
        "config": {"workload": (f"{len(states)} states, cantilever grids 30x15x15..80x40x20, "
                                "floors {1e-12,1e-3,1e-2}, FGMRES(100), tol 1e-6, guard"),
                   "accepted_fp32": acc32, "accepted_fp64": acc64, "solves": len(r32),
                   "plain_fp64_value": acc64 / w64,
                   "speedup_vs_plain_fp64": (acc32 / w32) / max(acc64 / w64, 1e-300),
                   "decisions_agree": sum(int(a["accepted"] == b["accepted"])
                                          for a, b in zip(r32, r64)),
                   "max_compliance_rel_diff_vs_fp64": max(crel) if crel else None,
                   "max_true_rel_residual_accepted": max(
                       [r["true_rel_residual"] for r in r32 if r["accepted"]] or [None]),
                   "mean_iters_fp32": float(np.mean([r["iters"] for r in r32])),
                   "mean_iters_fp64": float(np.mean([r["iters"] for r in r64])),
                   "parallelism": f"single GPU, {args.workers} concurrent states"},
    }
    if args.states_csv:
        import csv

        with open(args.states_csv, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=list(r32[0].keys()))
            w.writeheader()
            w.writerows(r32 + r64)
    print(json.dumps(line))


def coarsest_form(args):
    """Coarsest-solve form: the fixed-count Jacobi PCG of the coarsest level in
    its two-launch form (k_pcg_spmv + k_pcg_update per iteration, the default)
    and its persistent form (k_pcg_persistent, one launch per solve), timed
    through whole preconditioner applies of the canonical stack at --grid
    (use 200x100x50: its coarsest level has 6,825 free DOFs, above
    dense_chol_max, so it is solved by the PCG).  Elements are solid or void
    (1e-3 density, SIMP --penal) with probability 0.65 / 0.35 from
    --density-seed.  `value` is the persistent form's applies/s; the two
    forms differ only in the grouping of the PCG's dot-product partials."""
    import numpy as np
    import torch

    from gpu_fem import policy as gpol
    from gpu_fem.presets import cantilever

    dims = tuple(int(v) for v in args.grid.split("x"))
    n_e = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(args.density_seed)
    rho = np.where(rng.random(n_e) < 0.65, 1.0, 1e-3)
    E = torch.as_tensor(1e-9 + rho ** args.penal, device="cuda")
    res = {}
    for persistent in (False, True):
        comps = gpol._build_components(cantilever(*dims, load=args.load),
                                       coarse_pcg_persistent=persistent, **gpol.CANONICAL_STACK)
        gg = comps[7]
        gg.setup(E)
        n_coarsest = int(len(gg.level_free(gpol.CANONICAL_STACK["n_levels"] - 1)))
        rhs = torch.as_tensor(np.random.default_rng(11).standard_normal(len(comps[2])),
                              device="cuda")
        z = gg.apply(rhs)
        for _ in range(args.warmup):
            z = gg.apply(rhs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            z = gg.apply(rhs)
        torch.cuda.synchronize()
        res[persistent] = ((time.perf_counter() - t0) / args.steps * 1e3, z.cpu().numpy())
        print(f"{'persistent' if persistent else 'two-launch'}: {res[persistent][0]:.3f} ms "
              f"per apply", file=sys.stderr, flush=True)
        del comps, gg
    a, b = res[False][1], res[True][1]
    print(json.dumps({
        "metric": "preconditioner applies/s (coarsest level by persistent PCG)",
        "value": 1e3 / res[True][0], "unit": "applies/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res[True][0], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"canonical GMG preconditioner apply, cantilever "
                               f"{dims[0]}x{dims[1]}x{dims[2]}, coarsest {n_coarsest} free DOFs",
                   "two_launch_ms_per_apply": res[False][0],
                   "persistent_ms_per_apply": res[True][0],
                   "rel_diff_persistent_vs_two_launch":
                       float(np.abs(a - b).max() / np.abs(a).max())}}))


def precision_form(args):
    """Precision form: one filtered-sigmoid density state (seeded N(0,1) noise,
    cone filter radius 3, sigmoid sharpness 8, volume fraction 0.3) on the
    cantilever grid (default 200x100x50, 1,000,000 hex8), taken through the
    guarded policy with floors {1e-12 probe, 1e-3}, FGMRES restart 300, tol
    1e-6: once with the plain FP64 preconditioner and once with every
    multigrid level in FP32 (FP32-stored Galerkin stencils, FP32 smoothers;
    the outer FGMRES, the guard residual and the compliance stay FP64).
    Reports both forms' accepted solves/s and the FP32 result's compliance and
    solution error against the FP64 result (the FP64 form is the one held to
    the oracle), plus the fine-matvec rate at this grid.  A third run smooths
    the fine level with the bf16 matrix-core operator (levels 1.. FP32) and is
    reported beside them.  The forms run one at a time on one solver
    instance, on one hardware queue unless --hw-queues says otherwise (the
    co-residency rule then admits the 1M Gram-Schmidt sweep as one
    persistent launch per Arnoldi step)."""
    import numpy as np
    import torch

    from gpu_fem import policy as gpol
    from gpu_fem.batch import precision_stack
    from gpu_fem.density_field import filtered_sigmoid_density
    from gpu_fem.presets import cantilever

    dims = tuple(int(v) for v in args.grid.split("x"))
    n_e = dims[0] * dims[1] * dims[2]
    floors = tuple(float(v) for v in args.floors.split(","))
    progress(f"precision: density {dims}")
    rho = filtered_sigmoid_density(args.density_seed, dims, radius=3.0, sharpness=8.0,
                                   volfrac=0.3)
    spec = cantilever(*dims, load=args.load)
    kw = dict(penal=args.penal, baseline_rho_min=1e-12, raised_rho_mins=floors[1:],
              probe_iters=args.probe_iters, probe_r50=min(50, args.probe_iters),
              restart=args.restart, maxiter=args.maxiter, tol=1e-6, mgs_sweep=args.mgs_sweep,
              preset_name=f"cantilever_{dims[0]}x{dims[1]}x{dims[2]}_{args.load}")
    forms, xs, roof = {}, {}, {}
    for prec in ("fp64", "fp32", "bf16"):
        progress(f"precision {prec}: build")
        runner = gpol.GuardedFloorPolicy(spec, **kw, **precision_stack(prec))
        torch.cuda.synchronize()
        progress(f"precision {prec}: guarded policy")
        t0 = time.perf_counter()
        row = runner.run_state(rho_field=rho, keep_x=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        xs[prec] = row.pop("x")
        row["wall_s"] = wall
        forms[prec] = row
        if prec == "fp64":
            avg_us, bytes_alg, flops, gbs, tfs = matvec_roofline(runner)
            roof = {"avg_launch_us": avg_us, "algorithmic_bytes": bytes_alg, "flops": flops,
                    "hbm_GB_s": gbs, "fp64_tflops": tfs,
                    "launch_shape": runner.mf_op.launch_shape,
                    "fused_fp32_bf16": fused_matvec_rates(runner)}
        progress(f"precision {prec}: {wall:.1f} s, floor {row['recommended_rho_min']:.0e}, "
                 f"iters {row['solve_iters']}")
        del runner
        torch.cuda.empty_cache()
    f64, f32, fb = forms["fp64"], forms["fp32"], forms["bf16"]
    x64, x32 = xs["fp64"], xs["fp32"]
    sol_err = float(torch.linalg.vector_norm(x32 - x64) / torch.linalg.vector_norm(x64))
    sol_err_b = float(torch.linalg.vector_norm(xs["bf16"] - x64) / torch.linalg.vector_norm(x64))

    def brief(r):
        return {k: r[k] for k in ("trigger", "recommended_rho_min", "attempted_rho_mins",
                                  "probe_r50", "probe_r100", "solve_converged", "solve_iters",
                                  "solve_final_rel_residual", "compliance", "wall_s",
                                  "probe_setup_time_s", "probe_solve_time_s",
                                  "selected_setup_time_s", "selected_solve_time_s")}

    v32 = f32["solve_converged"] / f32["wall_s"]
    v64 = f64["solve_converged"] / f64["wall_s"]
    line = {
        "metric": "accepted guarded solves/s, precision form (FP32 multigrid) vs plain FP64",
        "value": v32, "unit": "accepted guarded solves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": f32["wall_s"] * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 outer / f32 preconditioner",
        "data": "synthetic (seeded filtered-sigmoid density)",
        "config": {
            "workload": (f"cantilever {dims[0]}x{dims[1]}x{dims[2]} ({n_e} hex8), filtered-sigmoid "
                         f"density seed {args.density_seed} (radius 3, sharpness 8, volume "
                         f"fraction 0.3), SIMP p={args.penal:g}, {args.load} load, floors "
                         f"{{{args.floors}}}, probe {args.probe_iters}, restart {args.restart}, "
                         f"maxiter {args.maxiter}, tol 1e-6"),
            "n_free": int(x64.numel()),
            "plain_fp64_value": v64,
            "speedup_vs_plain_fp64": v32 / max(v64, 1e-300),
            "fp32": brief(f32), "fp64": brief(f64),
            "floor_decisions_agree": int(f32["recommended_rho_min"] == f64["recommended_rho_min"]
                                         and f32["trigger"] == f64["trigger"]),
            "compliance_rel_err_vs_fp64": abs(f32["compliance"] - f64["compliance"])
            / abs(f64["compliance"]),
            "solution_rel_err_vs_fp64": sol_err,
            "fp32_true_rel_residual_fp64_recomputed": f32["solve_final_rel_residual"],
            "bf16_smoother": {
                "value": fb["solve_converged"] / fb["wall_s"],
                "speedup_vs_plain_fp64": (fb["solve_converged"] / fb["wall_s"]) / max(v64, 1e-300),
                "form": "fine level smoothed with the bf16 matrix-core operator, levels 1.. FP32",
                "run": brief(fb),
                "floor_decisions_agree": int(fb["recommended_rho_min"] == f64["recommended_rho_min"]
                                             and fb["trigger"] == f64["trigger"]),
                "compliance_rel_err_vs_fp64": abs(fb["compliance"] - f64["compliance"])
                / abs(f64["compliance"]),
                "solution_rel_err_vs_fp64": sol_err_b,
            },
            "fine_matvec_fp64": roof,
            "parallelism": "single GPU, one solver instance",
        },
    }
    print(json.dumps(line))


def learned_form(args):
    """Learned-floor form.  Training states (Bernoulli and filtered-sigmoid
    densities on the --train-grids) are labelled by the plain FP64 guarded
    policy -- or read from --labels, a CSV an earlier labelling run wrote --
    and the 3D CNN is trained on the GPU by the library's own BF16 kernels.
    Fresh-seed evaluation states on --grid then run through the plain FP64
    policy (probe + ladder + guard) on one solver instance and on concurrent
    instances, and through the learned policy variants (FP64 or FP32
    multigrid, optionally confidence-gated) on concurrent instances.
    ``value`` is the learned FP32-multigrid variant's accepted guarded
    solves/s; both plain rates, the speed-ups, the floor agreement with the
    plain ladder and the guard fallback rate are reported beside it."""
    import numpy as np
    import torch

    from gpu_fem import learned as L
    from gpu_fem.batch import precision_stack

    pk = dict(penal=args.penal)
    grids = [tuple(int(v) for v in g.split("x")) for g in args.train_grids.split(",") if g]
    train = L.sample_states(args.train_states, args.train_seed, grids)
    t_label = None
    labels = args.labels or ("" if args.relabel else L.DEFAULT_LABELS)
    if args.model:
        return learned_eval(args, L.FloorCNN.load(args.model), pk)
    if labels and os.path.exists(labels):
        progress(f"learned: reading labels {labels}")
        train, rows = L.read_labels(labels)
    else:
        progress(f"learned: labelling {len(train)} training states with the plain policy")
        t0 = time.perf_counter()
        rows = L.label_states(train, workers=args.workers, load=args.load, progress=progress,
                              **pk)
        t_label = time.perf_counter() - t0
        if args.labels_out:
            L.write_labels(args.labels_out, train, rows)
    progress("learned: pooling training inputs")
    L.attach_pooled(train, rows)
    X, feat, y_reg, y_cls = L.training_tensors(rows)
    model = L.FloorCNN(seed=0, trained_grids=sorted({tuple(r["dims"]) for r in rows}))
    progress(f"learned: training ({args.epochs} epochs)")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = model.fit(X, feat, y_reg, y_cls, epochs=args.epochs, batch=64, lr=3e-3, seed=0)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    cls_tr, _, conf_tr = model.predict_proba(X, feat)
    correct_tr = cls_tr == y_cls.cpu().numpy()
    train_acc = float(np.mean(correct_tr))
    if args.confidence == "auto":
        tau = L.confidence_for_accuracy(conf_tr, correct_tr, 0.99)
    else:
        tau = float(args.confidence)
    if args.theta_out:
        model.save(args.theta_out)

    dims = tuple(int(v) for v in args.grid.split("x"))
    kinds = tuple(k for k in args.eval_kinds.split(",") if k)
    vf = (args.eval_volfrac, args.eval_volfrac) if args.eval_volfrac else (0.2, 0.5)
    ev = L.sample_states(args.eval_states, args.eval_seed, [dims], kinds=kinds, volfrac_range=vf)
    dens = [L.density_of(s) for s in ev]
    workers = args.eval_workers or args.workers
    if args.eval_warmup:  # kernels, pools and autotuning of this grid
        progress("learned: warm-up")
        L.label_states(ev[:1], workers=1, load=args.load, densities=dens[:1], **pk)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rows = fn()
        torch.cuda.synchronize()
        return rows, time.perf_counter() - t0

    def acc(rs):
        return sum(int(r["solve_converged"]) for r in rs)

    progress(f"learned: plain FP64 policy, one instance, {len(ev)} evaluation states")
    plain_seq, w_seq = timed(lambda: L.label_states(ev, workers=1, load=args.load,
                                                    densities=dens, progress=progress, **pk))
    progress(f"learned: plain FP64 policy, {workers} concurrent instances")
    plain, w_plain = timed(lambda: L.label_states(ev, workers=workers, load=args.load,
                                                  densities=dens, progress=progress, **pk))
    variants = {"fp64": ("fp64", 0.0), "fp32": ("fp32", 0.0), "fp32_gated": ("fp32", tau),
                "fp64_gated": ("fp64", tau)}
    forms = {}
    for name in (v for v in args.learned_variants.split(",") if v):
        prec, gate = variants[name]
        progress(f"learned: learned policy, {prec} multigrid, confidence gate {gate:.3f}")
        got, wall = L.run_learned(model, ev, workers=workers, load=args.load, densities=dens,
                                  progress=progress, confidence_threshold=gate, **pk,
                                  **precision_stack(prec))
        forms[name] = (got, wall)

    v_seq = acc(plain_seq) / w_seq
    v_plain = acc(plain) / w_plain
    out = {}
    for name, (got, wall) in forms.items():
        c = L.compare(plain, got)
        c.update(value=acc(got) / wall, wall_s=wall,
                 speedup_vs_plain_fp64_one_instance=(acc(got) / wall) / v_seq,
                 speedup_vs_plain_fp64_concurrent=(acc(got) / wall) / v_plain,
                 mean_total_iters=float(np.mean([r["total_iters"] for r in got])))
        out[name] = c
    head = "fp32" if "fp32" in out else next(iter(out))
    w_head = forms[head][1]

    def iters(r):
        return len(r["probe_history"]) - 1 + sum(len(h) - 1 for _, h in r["solve_history"])

    line = {
        "metric": ("accepted guarded solves/s, learned floor + FP32 multigrid + concurrent "
                   "instances vs plain FP64"),
        "value": out[head]["value"], "unit": "accepted guarded solves/s", "n_gpus": 1,
        "steps": len(ev), "warmup": args.eval_warmup, "ms_per_step": w_head / len(ev) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 outer / f32 preconditioner; network bf16 (fp32 accumulation)",
        "data": "synthetic (seeded Bernoulli and filtered-sigmoid density states)",
        "config": {
            "workload": (f"{len(ev)} fresh-seed states on cantilever {dims[0]}x{dims[1]}x{dims[2]} "
                         f"({'/'.join(kinds)}; Bernoulli p in [0.1,0.5], filtered-sigmoid volume "
                         f"fraction in [{vf[0]},{vf[1]}]), SIMP p={args.penal:g}, {args.load} load, "
                         "ladder {1e-12,1e-3,1e-2}, restart 300, tol 1e-6"),
            "headline_variant": head,
            "training": {"states": len(train), "grids": [list(g) for g in grids],
                         "labelling_s": t_label, "train_s": t_train, "epochs": args.epochs,
                         "loss_first_last": [hist[0], hist[-1]], "train_class_accuracy": train_acc,
                         "labels": os.path.relpath(labels, ROOT) if t_label is None else None,
                         "confidence_gate": tau,
                         "train_fraction_above_gate": float(np.mean(conf_tr >= tau))},
            "plain_fp64_one_instance": {"value": v_seq, "wall_s": w_seq,
                                        "accepted": acc(plain_seq),
                                        "mean_total_iters": float(np.mean([iters(r) for r in plain_seq]))},
            "plain_fp64_concurrent": {"value": v_plain, "wall_s": w_plain, "accepted": acc(plain),
                                      "mean_total_iters": float(np.mean([iters(r) for r in plain]))},
            "plain_one_instance_equals_concurrent": all(
                a["recommended_rho_min"] == b["recommended_rho_min"]
                and a["solve_iters"] == b["solve_iters"] for a, b in zip(plain_seq, plain)),
            "learned": out,
            "parallelism": f"single GPU, {workers} concurrent solver instances",
        },
    }
    print(json.dumps(line))


def learned_eval(args, model, pk):
    """Learned form with a trained network (--model): fresh-seed states on
    --grid through the plain FP64 policy (one instance and concurrent
    instances) and through the learned policy with FP32 multigrid, ungated
    and gated by the network's per-grid calibrated confidence."""
    import numpy as np
    import torch

    from gpu_fem import learned as L
    from gpu_fem.batch import precision_stack

    dims = tuple(int(v) for v in args.grid.split("x"))
    kinds = tuple(k for k in args.eval_kinds.split(",") if k)
    vf = (args.eval_volfrac, args.eval_volfrac) if args.eval_volfrac else (0.2, 0.5)
    ev = L.sample_states(args.eval_states, args.eval_seed, [dims], kinds=kinds, volfrac_range=vf)
    dens = [L.density_of(s) for s in ev]
    workers = args.eval_workers or args.workers
    if args.eval_warmup:
        progress("learned: warm-up")
        L.label_states(ev[:1], workers=1, load=args.load, densities=dens[:1], **pk)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rows = fn()
        torch.cuda.synchronize()
        return rows, time.perf_counter() - t0

    def acc(rs):
        return sum(int(r["solve_converged"]) for r in rs)

    progress(f"learned: plain FP64 policy, one instance, {len(ev)} states on {dims}")
    plain_seq, w_seq = timed(lambda: L.label_states(ev, workers=1, load=args.load,
                                                    densities=dens, progress=progress, **pk))
    progress(f"learned: plain FP64 policy, {workers} concurrent instances")
    plain, w_plain = timed(lambda: L.label_states(ev, workers=workers, load=args.load,
                                                  densities=dens, progress=progress, **pk))
    v_seq, v_plain = acc(plain_seq) / w_seq, acc(plain) / w_plain
    out = {}
    for name, gate, prec in (("fp32_ungated", 0.0, "fp32"), ("fp32_gated", "auto", "fp32"),
                             ("fp64_gated", "auto", "fp64")):
        progress(f"learned: {name}")
        got, wall = L.run_learned(model, ev, workers=workers, load=args.load, densities=dens,
                                  progress=progress, confidence_threshold=gate, **pk,
                                  **precision_stack(prec))
        c = L.compare(plain, got)
        c.update(value=acc(got) / wall, wall_s=wall,
                 confidence_gate=model.gate_for(dims) if gate == "auto" else 0.0,
                 speedup_vs_plain_fp64_one_instance=(acc(got) / wall) / v_seq,
                 speedup_vs_plain_fp64_concurrent=(acc(got) / wall) / v_plain,
                 mean_total_iters=float(np.mean([r["total_iters"] for r in got])))
        out[name] = c
    print(json.dumps({
        "metric": ("accepted guarded solves/s, learned floor + FP32 multigrid + concurrent "
                   "instances vs plain FP64"),
        "value": out["fp32_gated"]["value"], "unit": "accepted guarded solves/s", "n_gpus": 1,
        "steps": len(ev), "warmup": args.eval_warmup,
        "ms_per_step": out["fp32_gated"]["wall_s"] / len(ev) * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 outer / f32 preconditioner; network bf16 (fp32 accumulation)",
        "data": "synthetic (seeded Bernoulli and filtered-sigmoid density states)",
        "config": {
            "workload": (f"{len(ev)} fresh-seed states on cantilever {dims[0]}x{dims[1]}x{dims[2]} "
                         f"({'/'.join(kinds)}), SIMP p={args.penal:g}, {args.load} load, ladder "
                         "{1e-12,1e-3,1e-2}, restart 300, tol 1e-6"),
            "model": args.model, "trained_grids": [list(g) for g in model.trained_grids or []],
            "in_training_range": model.in_range(dims),
            "plain_fp64_one_instance": {"value": v_seq, "wall_s": w_seq, "accepted": acc(plain_seq)},
            "plain_fp64_concurrent": {"value": v_plain, "wall_s": w_plain, "accepted": acc(plain)},
            "learned": out,
            "parallelism": f"single GPU, {workers} concurrent solver instances"}}))


def main():
    args = parse()
    if args.form == "precision" and not args.hw_queues:
        # one solver instance at a time: with one hardware queue the
        # co-residency rule admits the 1M-element Gram-Schmidt sweep (755
        # blocks) as one persistent launch per Arnoldi step
        args.hw_queues = 1
    if args.hw_queues:
        if not 1 <= args.hw_queues <= 32:
            raise SystemExit("--hw-queues must be in [1, 32]")
        # read by the HIP runtime at initialisation, which happens below
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(states shard across GPUs with no communication; "
                                   "multi-GPU runs are not part of this build)"}))
        sys.exit(2)
    import numpy as np
    import torch

    from gpu_fem import _lib
    from gpu_fem import policy as gpol

    if args.coarse_form != 1:
        _lib.call("gf_set_coarse_form", args.coarse_form)
    if args.ls_form != 1:
        _lib.call("gf_set_ls_form", args.ls_form)
    if args.form == "batched":
        batched_form(args)
        return
    if args.form == "precision":
        precision_form(args)
        return
    if args.form == "learned":
        learned_form(args)
        return
    if args.form == "coarsest":
        coarsest_form(args)
        return
    dims = tuple(int(v) for v in args.grid.split("x"))
    from gpu_fem.presets import cantilever

    spec = cantilever(*dims, load=args.load)
    kw = dict(penal=args.penal, raised_rho_mins=(1e-3, 1e-2), probe_iters=100, probe_r50=50,
              restart=300, maxiter=300, tol=1e-6, mgs_sweep=args.mgs_sweep,
              preset_name=f"cantilever_{dims[0]}x{dims[1]}x{dims[2]}_{args.load}")
    session = gpol.PolicySession(spec, workers=args.workers, **kw)
    runner = session.runners[0]
    gen = states(args)
    if args.roofline_only and args.roofline_kernel == "bf16":
        # the BF16 matrix-core form of the fused operator (one LDS-tiled
        # kernel), beside the FP32 form, for its MFMA-utilisation counters
        r = fused_matvec_rates(runner, n_launch=2000)
        src = torch.ones(8 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(50):
            dst.copy_(src)
        torch.cuda.synchronize()
        r.update(grid=list(dims), n_elem=runner.mf_op._n_elem,
                 copy_calibration_bytes={"read": 8 << 23, "write": 8 << 23})
        print(json.dumps(r))
        return
    if args.roofline_only and args.roofline_kernel == "stencil":
        rng = np.random.default_rng(3)
        n_e = dims[0] * dims[1] * dims[2]
        for r in session.runners:
            r.gmg.setup(torch.as_tensor(np.where(rng.random(n_e) < 0.35, 1.0, 1e-3),
                                        device="cuda"))
        r = stencil_roofline(session, spec, dict(gpol.CANONICAL_STACK, mgs_sweep=args.mgs_sweep),
                             n_launch=3000)
        src = torch.ones(8 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(50):
            dst.copy_(src)
        torch.cuda.synchronize()
        r.update(kernel="k_stencil_nodes (level 1, plain epilogue)",
                 copy_calibration_bytes={"read": 8 << 23, "write": 8 << 23})
        print(json.dumps(r))
        return
    if args.roofline_only:
        shapes = [int(v) for v in args.fine_shapes.split(",") if v.strip()] or [0]
        for shape in shapes:
            avg_us, bytes_alg, flops, gbs, tfs = matvec_roofline(runner, n_launch=2000,
                                                                 shape=shape)
            print(json.dumps({"kernel": "gf_matvec_f64", "avg_launch_us": avg_us,
                              "algorithmic_bytes": bytes_alg, "flops": flops, "GB_s": gbs,
                              "fp64_tflops": tfs,
                              "launch_shape": shape or runner.mf_op.launch_shape}))
        # calibration for HBM counters: a plain device copy of known size
        # (torch's copy kernel), 64 MiB read + 64 MiB written per launch
        src = torch.ones(8 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(50):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": {"read": 8 << 23, "write": 8 << 23}}))
        return
    n_e = dims[0] * dims[1] * dims[2]

    def batch(k):
        out = []
        for _ in range(k):
            seed, p = next(gen)
            mask = np.random.default_rng(seed).random(n_e) < p
            out.append({"solid": mask, "seed": seed, "probability": p})
        return out

    def finish(row):
        row["total_iters"] = (len(row["probe_history"]) - 1 +
                              sum(len(h) - 1 for _, h in row["solve_history"]))
        row["n_setups"] = 1 + row["n_floor_attempts"]
        return row

    # warm-up: the session's first state (sets the cached lambda for every
    # instance) and then one state per instance
    warm_states = batch(max(args.warmup, args.workers)) if args.warmup > 0 else []
    if warm_states:
        progress("policy: warm-up states")
        session.run(warm_states)
    progress(f"policy: {args.steps} timed states")
    torch.cuda.synchronize()
    timed = batch(args.steps)
    t0 = time.perf_counter()
    rows = [finish(r) for r in session.run(timed)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    accepted = sum(int(r["solve_converged"]) for r in rows)
    value = accepted / wall

    if args.no_extras:
        print(json.dumps({"metric": "accepted guarded solves/s (policy line only)",
                          "value": value, "unit": "accepted guarded solves/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": wall / args.steps * 1e3}))
        return
    # the plain policy on one solver instance, same states after the same
    # warm-up (instance 0 of the session: its rows must equal the concurrent ones)
    progress(f"policy: the same {args.steps} states on one instance")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(session.streams[0]):
        seq = [finish(runner.run_state(st["solid"], seed=st["seed"],
                                       probability=st["probability"], keep_x=False))
               for st in timed]
        torch.cuda.current_stream().synchronize()
    wall_seq = time.perf_counter() - t0
    value_seq = sum(int(r["solve_converged"]) for r in seq) / wall_seq
    seq_equal = all(a["recommended_rho_min"] == b["recommended_rho_min"]
                    and a["solve_iters"] == b["solve_iters"]
                    and a["solve_final_rel_residual"] == b["solve_final_rel_residual"]
                    for a, b in zip(rows, seq))
    learned = learned_on_states(args, spec, kw, dims, warm_states, timed, rows, wall, wall_seq)
    refp = reference_parametrisation(args)
    fine = matvec_roofline(runner)
    sten = stencil_roofline(session, spec, dict(gpol.CANONICAL_STACK, mgs_sweep=args.mgs_sweep))
    bf16 = fused_matvec_rates(runner)

    if args.states_csv:
        import csv

        keys = [k for k in gpol.SUMMARY_FIELDS if k in rows[0]] + ["total_iters", "n_setups"]
        with open(args.states_csv, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=keys, extrasaction="ignore")
            w.writeheader()
            w.writerows(rows)

    cpu = None
    if not args.no_cpu_baseline:
        progress("cpu baseline (oracle) sample")
        cpu = cpu_baseline(args, dims, rows)

    out = {
        "metric": "accepted guarded solves/s vs plain FP64 GMG-FGMRES @ true res<=tol; matvec GB/s",
        "value": value,
        "unit": "accepted guarded solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Bernoulli density states)",
        "config": {
            "workload": (f"plain FP64 guarded floor policy, cantilever {dims[0]}x{dims[1]}x{dims[2]} "
                         f"({dims[0] * dims[1] * dims[2]} hex8), SIMP p={args.penal:g}, "
                         f"{args.load} load, Bernoulli p in {{{args.probabilities}}}, ladder "
                         "{1e-12,1e-3,1e-2}, restart 300, tol 1e-6"),
            "grid": list(dims), "n_free": runner.mf_op._n_free,
            "states": [[r["seed"], r["solid_probability"]] for r in rows],
            "accepted": accepted,
            "mean_fgmres_iters_per_state": float(np.mean([r["total_iters"] for r in rows])),
            "triggers": [r["trigger"] for r in rows],
            "selected_floors": [r["recommended_rho_min"] for r in rows],
            "max_true_rel_residual": float(max(r["solve_final_rel_residual"] for r in rows)),
            "kept_floors_continuing_the_probe": sum(int(r.get("selected_continued_probe", 0))
                                                   for r in rows),
            "parallelism": f"single GPU, {args.workers} concurrent solver instances (one session)",
            "plain_fp64_one_instance": {
                "value": value_seq, "wall_s": wall_seq,
                "rows_equal_concurrent": seq_equal,
                "concurrent_speedup": value / value_seq},
            "learned": learned,
            "reference_parametrisation": refp,
        },
        "roofline": {
            "kernel": ("k_stencil_nodes<double, plain epilogue> (27-point 3x3-block coarse "
                       "operator, level 1 of the hierarchy; gf_gmg_level_matvec)"),
            "bound": "hbm",
            "achieved": sten["GB_s"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": sten["GB_s"] / HBM_PEAK_GBS,
            "traffic": STENCIL_TRAFFIC_BYTES,
            "note": ("the policy's largest kernel family (k_stencil_nodes 27.2 % of the "
                     "summed kernel time, the fine operator 23.9 %, the Gram-Schmidt sweep "
                     "21.7 %, profiles/policy_state_80x40x20_fp64_kernel_stats.md), with an "
                     "HBM roof: 0.24 flop/B; timed round-robin over set-up hierarchies whose "
                     "stencils together exceed the 256 MB Infinity Cache; the fine operator's "
                     "FP64-compute roofline is reported beside it"),
            **sten,
        },
        "fine_operator": {
            "kernel": (f"{SHAPE_KERNELS.get(runner.mf_op.launch_shape, '?')} FP64, "
                       "plain epilogue (gf_matvec_f64)"),
            "bound": "fp64 vector ALU (19.3 flop/B, above the FP64 ridge of 9.8 flop/B)",
            "achieved_tflops": fine[4], "peak_tflops": FP64_PEAK_TFLOPS,
            "frac": fine[4] / FP64_PEAK_TFLOPS, "avg_launch_us": fine[0],
            "algorithmic_bytes": fine[1], "flops": fine[2], "hbm_GB_s": fine[3],
            "traffic": TRAFFIC_BYTES,
            "launch_shape": runner.mf_op.launch_shape,
            "autotune_us": runner.mf_op.launch_timings_us,
            "fused_matvec_forms": bf16,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
