"""Host-side checks that need no GPU: the C ABI library loads and exports every
symbol include/gfhip.h declares; the product's host set-up (stiffness, DOF
table, presets, boundary conditions, classification, connectivity) equals the
oracle's."""

import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "gfhip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gf_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from gpu_fem import _lib

    lib = _lib.load()
    names = _declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature"
    assert lib.gf_version() >= 10000


def test_params_struct_layout_matches_header():
    import ctypes as C

    from gpu_fem import _lib

    lib = _lib.load()
    p = _lib.gf_gmg_params()
    lib.gf_gmg_default_params(C.byref(p))
    assert p.n_levels == 4 and p.coarse_smooth_iters == 2 and p.dense_chol_max == 5000
    assert abs(p.cheb_lower_frac - 1 / 30) < 1e-15 and p.coarse_pcg_iters == 80
    assert list(p.level_prec) == [-1] * 8 and p.root_enrichment_weight_floor == 1e-6
    assert p.mgs_sweep == 1 and p.coarse_pcg_persistent == 1 and p.graph_apply == 1


def test_library_reports_argument_errors_without_gpu():
    import ctypes as C

    from gpu_fem import _lib

    lib = _lib.load()
    h = C.c_void_p()
    free = np.arange(10, dtype=np.int32)
    ke = np.zeros(576)
    rc = lib.gf_ctx_create(C.byref(h), 2, 2, 2, ke.ctypes.data, free.ctypes.data, 10, 12345)
    assert rc == -2
    assert b"ndof" in lib.gf_last_error()
    # process-wide kernel-form switches reject unknown forms (no device work)
    assert lib.gf_set_sweep_form(3) == -2
    assert b"gf_set_sweep_form" in lib.gf_last_error()
    assert lib.gf_set_ls_form(2) == -2
    assert b"gf_set_ls_form" in lib.gf_last_error()
    assert lib.gf_set_coarse_form(7) == -2
    assert b"gf_set_coarse_form" in lib.gf_last_error()


def test_host_setup_equals_oracle():
    from gpu_fem import bc_generator, presets
    from gpu_fem.pub_simp_solver import KE_UNIT_3D, _edof_table_3d, grid_from_edof
    from oracle import fem

    assert np.allclose(KE_UNIT_3D, fem.KE_UNIT_3D, rtol=0, atol=2e-16)
    for dims in [(3, 2, 2), (8, 4, 5)]:
        ed = _edof_table_3d(*dims)
        assert np.array_equal(ed, fem.edof_hex8(*dims))
        assert grid_from_edof(ed) == dims
    with pytest.raises(ValueError):
        grid_from_edof(_edof_table_3d(3, 2, 2)[::-1])
    for name in ["cantilever_gpu_medium", "bridge_gpu_medium", "cantilever_3d", "bracket_3d",
                 "bracket_gpu_medium"]:
        a = bc_generator.generate_bc(presets.get_preset(name))
        b = fem.generate_bc(fem.get_preset(name))
        assert np.array_equal(a.free_dofs, b.free_dofs)
        assert np.array_equal(a.F, b.F)
    # bracket (presets.py:363-369): the y = Ly face is clamped, and the point
    # load (fx, fy) = (0.3, -0.7) lands on node (nelx, 0, nelz/2)
    nx, ny, nz = 30, 60, 15
    b = fem.generate_bc(fem.get_preset("bracket_gpu_medium"))
    assert b.ndof - len(b.free_dofs) == 3 * (nx + 1) * (nz + 1)
    node = (nx * (ny + 1) + 0) * (nz + 1) + nz // 2 + (nz % 2)
    nz_idx = np.nonzero(b.F)[0]
    assert set(nz_idx) <= {3 * node, 3 * node + 1, 3 * (node - 1), 3 * (node - 1) + 1}
    assert np.isclose(b.F.sum(), -0.4) and np.isclose(b.F[nz_idx[0]], 0.3)
    a = bc_generator.generate_bc(presets.cantilever(40, 20, 20, load="edge"))
    b = fem.generate_bc(fem.cantilever_spec(40, 20, 20, load="edge"))
    assert np.array_equal(a.F, b.F) and np.array_equal(a.free_dofs, b.free_dofs)


def test_classify_and_connectivity_equal_oracle():
    from gpu_fem.policy import _classify, _connectivity
    from oracle import policy as opol

    thr = dict(high_residual_threshold=1e-2, plateau_residual_threshold=1e-4,
               plateau_ratio_threshold=0.6)
    rng = np.random.default_rng(0)
    for _ in range(500):
        r50, r100 = 10 ** rng.uniform(-8, 0), 10 ** rng.uniform(-8, 0)
        assert _classify(r50, r100, **thr) == opol.classify(r50, r100, **thr)
    m = opol.bernoulli_mask(5, (20, 10, 10), 0.3)
    assert _connectivity(m) == opol.connectivity(m)


def test_filtered_sigmoid_density_matches_oracle():
    from gpu_fem.density_field import filtered_sigmoid_density
    from oracle.policy import filtered_sigmoid_density as ofsd

    for seed, dims in ((0, (12, 6, 6)), (3, (20, 10, 5))):
        a = filtered_sigmoid_density(seed, dims)
        b = ofsd(seed, dims)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-14)
        assert abs(a.mean() - 0.3) < 1e-9 and a.min() > 0.0 and a.max() < 1.0


def test_density_csv_schema_and_history_rows(tmp_path):
    import csv

    from gpu_fem import density_field as dfm

    row = {k: i for i, k in enumerate(dfm.DENSITY_SUMMARY_FIELDS)}
    row.update(probe_history=[1.0, 0.5, 0.1], solve_history=[(1e-3, [1.0, 1e-7])],
               extra_key="ignored")
    p = tmp_path / "summary.csv"
    dfm.write_csv(str(p), [row], dfm.DENSITY_SUMMARY_FIELDS)
    with open(p) as f:
        r = list(csv.reader(f))
    assert r[0] == dfm.DENSITY_SUMMARY_FIELDS and len(r) == 2
    h = dfm.history_rows(row, {"density_name": "d", "density_path": "x.npy"}, 1e-12)
    assert [x["phase"] for x in h] == ["probe"] * 3 + ["recommended_solve"] * 2
    assert [x["iter"] for x in h] == [0, 1, 2, 0, 1]
    assert h[3]["rho_min"] == 1e-3 and h[0]["rho_min"] == 1e-12
    q = tmp_path / "hist.csv"
    dfm.write_csv(str(q), h, dfm.DENSITY_HISTORY_FIELDS)
    with open(q) as f:
        assert next(csv.reader(f)) == dfm.DENSITY_HISTORY_FIELDS


# ---- learned-floor form: host logic ----
def test_learned_ladder_rules():
    from gpu_fem.learned import learned_ladder

    floors = (1e-12, 1e-3, 1e-2)
    verdicts = {1e-12: False, 1e-3: True, 1e-2: True}
    assert learned_ladder(1, floors, verdicts.get) == ([1e-3], 1, False)
    assert learned_ladder(0, floors, verdicts.get) == ([1e-12, 1e-3], 1, True)
    assert learned_ladder(2, floors, verdicts.get) == ([1e-2], 2, False)
    assert learned_ladder(0, floors, lambda r: False) == ([1e-12, 1e-3, 1e-2], None, True)
    with pytest.raises(ValueError):
        learned_ladder(3, floors, verdicts.get)


def test_learned_targets_and_classes():
    from gpu_fem.learned import FLOORS, floor_class, grid_features, r50_from_target, r50_target

    r = np.array([1e-9, 3.2e-5, 0.7])
    assert np.allclose(r50_from_target(r50_target(r)), r, rtol=1e-12)
    # breakdown (0), divergence (inf) and NaN probes are clipped
    assert np.allclose(r50_target([0.0, np.inf, np.nan, 1e-20]),
                       [(-12 + 5) / 3, (0.5 + 5) / 3, (0.5 + 5) / 3, (-12 + 5) / 3])
    assert [floor_class(f) for f in FLOORS] == [0, 1, 2]
    assert floor_class(1.0000001e-3) == 1
    assert np.allclose(grid_features((32, 64, 16)), [0.0, 1.0, -1.0])


def test_learned_init_matches_oracle_initialiser():
    from gpu_fem.learned import init_theta
    from oracle import cnn

    assert np.array_equal(init_theta(4), cnn.init_params(4))


def test_learned_state_sampling_is_seeded():
    from gpu_fem.learned import density_of, sample_states

    grids = [(40, 20, 20), (30, 15, 15)]
    a = sample_states(6, 3, grids)
    assert a == sample_states(6, 3, grids)
    assert [s["kind"] for s in a] == ["bernoulli", "sigmoid"] * 3
    d0, d1 = density_of(a[0]), density_of(a[1])
    assert set(np.unique(d0)) <= {0.0, 1.0}
    assert abs(d1.mean() - a[1]["volfrac"]) < 1e-6 and 0.0 < d1.min() and d1.max() <= 1.0
    assert 0.05 < np.mean((d1 > 0.05) & (d1 < 0.95)) < 0.95  # intermediate densities


def test_learned_compare_counts():
    from gpu_fem.learned import compare

    plain = [{"recommended_rho_min": 1e-12, "solve_converged": 1, "probe_r50": 1e-6},
             {"recommended_rho_min": 1e-3, "solve_converged": 1, "probe_r50": 1e-2}]
    learned = [{"recommended_rho_min": 1e-12, "predicted_class": 0, "guard_fallback": 0,
                "solve_converged": 1, "solve_final_rel_residual": 5e-7, "predicted_r50": 1e-6},
               {"recommended_rho_min": 1e-2, "predicted_class": 2, "guard_fallback": 0,
                "solve_converged": 1, "solve_final_rel_residual": 9e-7, "predicted_r50": 1e-1}]
    c = compare(plain, learned)
    assert c["floor_agreement"] == 0.5 and c["proposal_agreement"] == 0.5
    assert c["fallback_rate"] == 0.0 and c["max_accepted_true_rel_residual"] == 9e-7
    assert abs(c["r50_log10_rmse"] - np.sqrt(0.5)) < 1e-12


def test_learned_label_csv_round_trip(tmp_path):
    """Labels written by the labelling run read back as the same state
    descriptors (so the densities regenerate) and the same policy outcome,
    with floats exact."""
    from gpu_fem.learned import density_of, read_labels, sample_states, write_labels

    states = sample_states(4, 9, [(12, 6, 6), (10, 5, 5)])
    rows = [{"trigger": "keep" if i % 2 else "high_r50", "probe_r50": 10.0 ** (-3 - i / 7),
             "probe_r100": 0.1 / (i + 3), "recommended_rho_min": (1e-12, 1e-3)[i % 2],
             "n_floor_attempts": 1 + i % 2, "solve_converged": 1, "solve_iters": 30 + i,
             "recorded_policy_time_s": 0.123 * (i + 1)} for i in range(4)]
    p = str(tmp_path / "labels.csv")
    write_labels(p, states, rows)
    st2, rows2 = read_labels(p)
    assert st2 == states
    for a, b in zip(rows, rows2):
        for k, v in a.items():
            assert b[k] == v, k
    for a, b in zip(states, st2):
        assert np.array_equal(density_of(a), density_of(b))


def test_confidence_for_accuracy():
    from gpu_fem.learned import confidence_for_accuracy

    conf = np.array([0.99, 0.95, 0.9, 0.8, 0.7, 0.6, 0.5])
    ok = np.array([1, 1, 1, 0, 1, 0, 1], dtype=bool)
    # prefix accuracies by falling confidence: 1, 1, 1, 3/4, 4/5, 4/6, 5/7
    assert confidence_for_accuracy(conf, ok, 0.99) == 0.9
    assert confidence_for_accuracy(conf, ok, 0.8) == 0.7
    assert confidence_for_accuracy(conf, ok, 0.7) == 0.5
    assert confidence_for_accuracy(conf, np.zeros(7, bool), 0.5) == 1.0


def test_device_lanczos_recurrence_on_the_host():
    """The package's Lanczos recurrence (``lanczos_kappa``, run on device
    vectors through the library's dot / nrm2 / axpby in the GPU path) with
    NumPy stand-ins for those callables: exact max/min on a diagonal operator
    with n <= steps (breakdown path), and the oracle's recurrence to rounding
    on a longer spectrum."""
    from gpu_fem.multigrid_v4 import lanczos_kappa
    from oracle.gmg import lanczos_kappa as oracle_lanczos

    def axpby(a, x, b, y):
        y *= b
        y += a * x
        return y

    kw = dict(dot_fn=lambda a, b: float(a @ b), nrm_fn=lambda a: float(np.linalg.norm(a)),
              axpby_fn=axpby, zeros_like=np.zeros_like)
    lam = np.array([0.5, 1.0, 2.0, 3.0, 7.5, 10.0])
    v = np.random.default_rng(1).standard_normal(len(lam))
    v /= np.linalg.norm(v)
    assert abs(lanczos_kappa(lambda u: lam * u, v.copy(), 30, **kw) - 20.0) <= 1e-10 * 20.0
    lam = np.linspace(1.0, 100.0, 200)
    v = np.random.default_rng(2).standard_normal(len(lam))
    v /= np.linalg.norm(v)
    for m in (5, 12, 25):
        a = lanczos_kappa(lambda u: lam * u, v.copy(), m, **kw)
        b = oracle_lanczos(lambda u: lam * u, v.copy(), m)
        assert abs(a - b) <= 1e-9 * b, m


def test_trace_ratio_rules():
    """_safe_ratio, _merge_notes and _alignment follow multigrid_v4.py:840-862
    and 934-936 (zero denominators, empty notes)."""
    from gpu_fem.multigrid_v4 import _alignment, _merge_notes, _safe_ratio

    assert _safe_ratio(1.0, 2.0) == 0.5
    assert _safe_ratio(0.0, 0.0) == 0.0
    assert _safe_ratio(1.0, 0.0) == float("inf")
    assert _merge_notes("", "a", "", "b") == "a|b"
    assert _merge_notes("", "") == ""
    assert _alignment(0.0, 1.0, 5.0) == 0.0
    assert _alignment(2.0, 3.0, 6.0) == 1.0


def test_learned_grid_range_gate():
    """A state whose grid lies outside the per-axis range of the training
    grids is not proposed for (grid features would be extrapolated)."""
    from gpu_fem.learned import DEFAULT_LABELS, in_grid_range, read_labels

    grids = [(40, 20, 20), (64, 32, 16), (80, 40, 20)]
    assert in_grid_range(grids, (80, 40, 20)) and in_grid_range(grids, (50, 30, 18))
    assert not in_grid_range(grids, (160, 80, 40))
    assert not in_grid_range(grids, (24, 12, 12))
    assert not in_grid_range(grids, (81, 40, 20))
    assert in_grid_range(None, (200, 100, 50)) and in_grid_range([], (1, 1, 1))
    shipped = sorted({tuple(s["dims"]) for s in read_labels(DEFAULT_LABELS)[0]})
    assert shipped and all(in_grid_range(shipped, g) for g in shipped)


def test_shipped_network_file():
    """The shipped network: parameter count of the architecture, training
    grids spanning 16k..1M elements, a confidence gate per training grid, and
    label files for every training grid."""
    import numpy as np
    from gpu_fem.learned import DEFAULT_MODEL, LABEL_FILES, read_labels

    z = np.load(DEFAULT_MODEL, allow_pickle=False)
    assert z["theta"].shape == (15588,) and z["theta"].dtype == np.float32
    grids = {tuple(int(v) for v in g) for g in z["trained_grids"]}
    assert (40, 20, 20) in grids and (200, 100, 50) in grids
    assert {tuple(int(v) for v in g) for g in z["gate_grids"]} == grids
    assert np.all((z["gate_tau"] > 0) & (z["gate_tau"] <= 1))
    labelled = {tuple(s["dims"]) for f in LABEL_FILES for s in read_labels(f)[0]}
    assert labelled == grids


def test_bench_refuses_more_than_one_gpu_and_bad_queue_counts():
    """bench.py runs on one GPU: --gpus above 1 prints one JSON line that
    says so and exits 2 before touching a device; a hardware-queue count
    outside [1, 32] is refused the same way."""
    import json
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    bench = os.path.join(root, "bench.py")
    r = subprocess.run([sys.executable, bench, "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]
    r = subprocess.run([sys.executable, bench, "--hw-queues", "40"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "--hw-queues must be in [1, 32]" in r.stderr
