"""Pins for the oracle's policy layer against the reference's own result tables.

* Bernoulli masks + 6-neighbour connectivity reproduce the integer columns
  solid_count / n_components / largest_component /
  support_to_load_patch_components of the held-out cantilever (80x40x20) and
  bridge (90x30x15) suites exactly: this pins the PCG64 stream, the element
  ordering of the density field and the BFS.
* ``classify`` reproduces the ``trigger`` column of every summary row from that
  row's own probe_r50 / probe_r100 columns.
* ``run_policy``'s bookkeeping reproduces the attempted-floor / fallback /
  false-keep columns of every row from its trigger and convergence columns,
  and scripted solvers whose projected residual lies are rejected by the guard.
"""

import csv
import os

import numpy as np

from oracle import policy as opol

DIMS = {"cantilever": (80, 40, 20), "bridge": (90, 30, 15)}


def _rows(golden_dir, name):
    with open(os.path.join(golden_dir, name)) as f:
        return list(csv.DictReader(f))


def _suite(golden_dir):
    out = []
    for name, kind in (("heldout_cantilever_guarded.csv", "cantilever"),
                       ("heldout_bridge_guarded.csv", "bridge"),
                       ("heldout_cantilever_strict.csv", "cantilever")):
        for r in _rows(golden_dir, name):
            out.append((kind, r))
    return out


def _check_connectivity(kind, r):
    mask = opol.bernoulli_mask(int(r["seed"]), DIMS[kind], float(r["solid_probability"]))
    c = opol.connectivity(mask)
    for k in ("solid_count", "n_components", "largest_component",
              "support_to_load_patch_components"):
        assert c[k] == int(r[k]), (kind, r["seed"], r["solid_probability"], k, c[k], r[k])


def test_connectivity_all_rows_match_reference_tables(golden_dir):
    for kind, r in _suite(golden_dir):
        _check_connectivity(kind, r)


def test_known_mask_counts():
    # the reference's quoted example: seed 47, p 0.35 on 80x40x20
    c = opol.connectivity(opol.bernoulli_mask(47, (80, 40, 20), 0.35))
    assert (c["solid_count"], c["n_components"], c["largest_component"]) == (22266, 2836, 13553)


def test_connectivity_brute_force_small():
    # independent union-find over the same 6-neighbour graph
    rng = np.random.default_rng(5)
    for _ in range(20):
        m = rng.random((5, 6, 7)) < 0.45
        idx = {p: i for i, p in enumerate(zip(*np.nonzero(m)))}
        parent = list(range(len(idx)))

        def find(a):
            while parent[a] != a:
                parent[a] = parent[parent[a]]
                a = parent[a]
            return a

        for (x, y, z), i in idx.items():
            for d in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                q = (x + d[0], y + d[1], z + d[2])
                if q in idx:
                    parent[find(i)] = find(idx[q])
        roots = [find(i) for i in range(len(idx))]
        sizes = np.bincount(roots) if roots else np.zeros(1, int)
        c = opol.connectivity(m)
        assert c["n_components"] == len(set(roots))
        assert c["largest_component"] == (sizes.max() if roots else 0)
        assert c["solid_count"] == len(idx)


def test_classify_reproduces_trigger_column(golden_dir):
    n = 0
    for _, r in _suite(golden_dir):
        raise_floor, trig = opol.classify(float(r["probe_r50"]), float(r["probe_r100"]))
        assert trig == r["trigger"], (r["seed"], r["solid_probability"], trig, r["trigger"])
        assert int(raise_floor) == int(r["predicted_raise_floor"])
        n += 1
    assert n == 72 + 30 + 72


def test_classify_truth_table_edges():
    assert opol.classify(1e-2, 0.0) == (True, "high_r50")          # r50 threshold inclusive
    assert opol.classify(0.99e-2, 1e-4)[1] == "keep"                # ratio 0.0101 < 0.6
    assert opol.classify(1e-4, 0.6e-4) == (False, "keep")           # r100 below plateau level
    assert opol.classify(1e-3, 0.6e-3) == (True, "plateau")         # ratio exactly 0.6
    assert opol.classify(1e-3, 0.59e-3) == (False, "keep")


def test_ladder_bookkeeping_columns(golden_dir):
    """attempted floors, fallback and false-keep follow from trigger and the
    per-floor convergence outcome (the guard) exactly as the rows record."""
    for _, r in _suite(golden_dir):
        attempted = [float(v) for v in r["attempted_rho_mins"].split(";")]
        keep = r["trigger"] == "keep"
        if "post_solve_fallback_triggered" not in r:
            # pre-guard replay (strict table): no fallback ladder existed
            assert attempted == ([1e-12] if keep else attempted[:1] + attempted[1:])
            assert float(r["recommended_rho_min"]) == attempted[-1]
            continue
        fb = int(r["post_solve_fallback_triggered"])
        if keep:
            assert attempted[0] == 1e-12
            assert fb == int(len(attempted) > 1)
            assert int(r["detector_false_keep"]) == fb
        else:
            assert attempted[0] == 1e-3 and fb == 0
        assert float(r["recommended_rho_min"]) == attempted[-1]


def _fake(n, script):
    """A solver whose reported history and returned x are scripted per floor:
    script[rho] = (true_solution_ok, projected_history)."""
    F = np.ones(n)

    def A_factory(E):
        return lambda x: x

    def solve(E, mx, hist):
        # the modulus of an all-void state is rho_min (+ rho_min**penal)
        rho = min(script, key=lambda k: abs(np.log(k) - np.log(float(E.min()))))
        ok, h = script[rho]
        hist.extend(h[: mx + 1])
        x = F.copy() if ok else 0.5 * F
        return x, len(h) - 1, True
    return F, solve, A_factory


def test_guard_rejects_lying_solver_and_ladder_advances():
    n = 16
    solid = np.zeros(n, dtype=bool)
    conv = [1.0] + [1e-7] * 100
    script = {1e-12: (False, conv), 1e-3: (False, conv), 1e-2: (True, conv)}
    F, solve, Af = _fake(n, script)
    row = opol.run_policy(solid, solve, F, Af)
    assert row["trigger"] == "keep"
    assert row["solver_reported_converged"] == 1
    assert row["post_solve_fallback_triggered"] == 1 and row["detector_false_keep"] == 1
    assert row["attempted_rho_mins"] == [1e-12, 1e-3, 1e-2]
    assert row["recommended_rho_min"] == 1e-2 and row["solve_converged"] == 1


def test_raised_ladder_stops_at_first_guarded_floor():
    n = 8
    solid = np.zeros(n, dtype=bool)
    bad = [1.0] + [0.5] * 100                 # r50 = 0.5 -> high_r50
    script = {1e-12: (False, bad), 1e-3: (True, [1.0, 1e-7]), 1e-2: (True, [1.0, 1e-7])}
    F, solve, Af = _fake(n, script)
    row = opol.run_policy(solid, solve, F, Af)
    assert row["trigger"] == "high_r50" and row["attempted_rho_mins"] == [1e-3]
    assert row["solve_iters"] == 1 and row["post_solve_fallback_triggered"] == 0


def test_short_probe_history_clamps_r50_r100():
    n = 8
    solid = np.zeros(n, dtype=bool)
    short = [1.0, 0.3, 1e-7]                  # converged after 2 iterations
    script = {1e-12: (True, short), 1e-3: (True, short), 1e-2: (True, short)}
    F, solve, Af = _fake(n, script)
    row = opol.run_policy(solid, solve, F, Af)
    assert row["probe_r50"] == 1e-7 and row["probe_r100"] == 1e-7
    assert row["trigger"] == "keep"
