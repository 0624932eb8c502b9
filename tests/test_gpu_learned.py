"""Learned floor predictor on the GPU: the CNN kernels (pool, forward, BF16
gradient, Adam) against the oracle (oracle/cnn.py), training determinism,
and the learned policy path against the plain policy."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _t(a, dtype=None):
    import torch

    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=dtype)


def _batch(B, seed):
    rng = np.random.default_rng(seed)
    X = rng.random((B, 16, 8, 8)).astype(np.float32)
    feat = rng.uniform(-1, 2, size=(B, 3)).astype(np.float32)
    y_reg = rng.normal(size=B).astype(np.float32)
    y_cls = rng.integers(0, 3, size=B).astype(np.int32)
    return X, feat, y_reg, y_cls


def _theta(seed):
    from oracle import cnn

    th = cnn.init_params(seed)
    return (th + 0.05 * np.random.default_rng(seed + 100).normal(size=th.shape)).astype(np.float32)


@pytest.mark.parametrize("dims", [(40, 20, 20), (33, 9, 17), (160, 80, 40), (7, 3, 5)])
def test_pool_matches_oracle(dims):
    from gpu_fem.learned import pool_density
    from oracle import cnn

    rho = np.random.default_rng(sum(dims)).random(int(np.prod(dims)))
    X = pool_density(rho, dims).cpu().numpy()
    ref = cnn.adaptive_pool(rho, dims)
    # float32 rounding of a double mean
    assert np.abs(X - ref).max() <= 1e-7 * max(1.0, np.abs(ref).max())


def test_forward_matches_oracle():
    from gpu_fem.learned import FloorCNN
    from oracle import cnn

    th = _theta(1)
    X, feat, _, _ = _batch(6, 2)
    out = FloorCNN(theta=th).forward(_t(X), _t(feat)).cpu().numpy()
    ref, _ = cnn.forward(th, X, feat, bf16=True)
    err = np.abs(out - ref).max() / np.abs(ref).max()
    print("forward rel err", err)
    # FP32 accumulation against float64 sums of the same bf16 operands, plus
    # rare bf16 rounding-boundary flips of intermediate activations
    assert err <= 2e-4


def test_gradient_matches_oracle():
    from gpu_fem.learned import FloorCNN
    from oracle import cnn

    th = _theta(3)
    X, feat, y_reg, y_cls = _batch(5, 4)
    model = FloorCNN(theta=th)
    loss, g, out = model.grad(_t(X), _t(feat), _t(y_reg), _t(y_cls), 0.8)
    loss, g = float(loss.item()), g.cpu().numpy()
    l_ref, g_ref, _ = cnn.loss_and_grad(th, X, feat, y_reg, y_cls, 0.8, bf16=True)
    assert abs(loss - l_ref) <= 1e-4 * abs(l_ref)
    for name, (o, shape) in cnn.OFFSETS.items():
        n = int(np.prod(shape))
        a, b = g[o:o + n], g_ref[o:o + n]
        scale = np.abs(b).max()
        assert scale > 0, name
        d = np.abs(a - b) / scale
        print(name, "max", d.max(), "median", np.median(d))
        # every entry: within two bf16 half-ulps of the block scale (an
        # operand whose FP32 value sits on a bf16 rounding boundary may round
        # the other way than in the float64 oracle); almost every entry: FP32
        # accumulation accuracy
        assert d.max() <= 8e-3, name
        assert np.mean(d <= 1e-4) >= 0.9, name


def test_adam_matches_oracle():
    from gpu_fem.learned import FloorCNN
    from oracle import cnn

    rng = np.random.default_rng(5)
    th = rng.normal(size=cnn.N_PARAMS).astype(np.float32)
    model = FloorCNN(theta=th)
    t64, m64, v64 = th.astype(np.float64), np.zeros(cnn.N_PARAMS), np.zeros(cnn.N_PARAMS)
    for k in range(1, 4):
        g = rng.normal(size=cnn.N_PARAMS).astype(np.float32)
        model.adam(_t(g), 1e-2)
        t64, m64, v64 = cnn.adam_step(t64, m64, v64, g, k, 1e-2)
    got = model.theta_host()
    assert model.step == 3
    assert np.abs(got - t64).max() <= 1e-5 * max(1.0, np.abs(t64).max())


def test_training_is_deterministic_and_learns_a_separable_task():
    """Labels that are a function of the pooled density (class by mean
    density, regression target = 2 * mean - 1): the loss must fall and the
    classes be recovered; two runs with the same seed give identical
    parameters."""
    from gpu_fem.learned import FloorCNN

    rng = np.random.default_rng(11)
    n = 192
    level = rng.uniform(0.05, 0.95, size=n)
    X = np.clip(level[:, None, None, None] + 0.05 * rng.normal(size=(n, 16, 8, 8)), 0, 1)
    X = X.astype(np.float32)
    feat = np.zeros((n, 3), np.float32)
    y_cls = np.digitize(level, [0.35, 0.65]).astype(np.int32)
    y_reg = (2 * level - 1).astype(np.float32)
    thetas = []
    for _ in range(2):
        m = FloorCNN(seed=0)
        hist = m.fit(_t(X), _t(feat), _t(y_reg), _t(y_cls), epochs=40, batch=32, lr=1e-2,
                     seed=3)
        thetas.append(m.theta_host())
    print("loss", hist[0], hist[-1])
    assert hist[-1] < 0.3 * hist[0]
    cls, _ = m.predict(_t(X), _t(feat))
    assert np.mean(cls == y_cls) >= 0.9
    assert np.array_equal(thetas[0], thetas[1])


def test_learned_policy_matches_plain_selected_solve_when_proposal_agrees():
    """Proposing exactly the plain policy's accepted floor reproduces its
    selected solve: same iteration count and compliance (one session each,
    same first set-up moduli, so the cached spectral estimate is equal)."""
    from gpu_fem.learned import (FLOORS, LearnedFloorPolicy, POLICY_DEFAULTS, floor_class,
                                 label_states, sample_states)
    from gpu_fem.policy import PolicySession
    from gpu_fem.presets import cantilever

    states = sample_states(4, 7, [(24, 12, 12)], kinds=("bernoulli",), p_range=(0.3, 0.5))
    for st in states:
        row = label_states([st], workers=1)[0]  # a session of its own
        k = floor_class(row["recommended_rho_min"])
        if k != 0:
            continue  # the plain selected solve follows a probe at another floor
        session = PolicySession(cantilever(*st["dims"], load="edge"), workers=1,
                                runner_cls=LearnedFloorPolicy, **POLICY_DEFAULTS)
        rho = (np.random.default_rng(st["seed"]).random(int(np.prod(st["dims"])))
               < st["probability"])
        got = session.run([{"solid": rho, "extra": {"predicted_class": k}}])[0]
        assert got["guard_fallback"] == 0
        assert got["solve_iters"] == row["solve_iters"]
        assert got["recommended_rho_min"] == FLOORS[k]
        assert abs(got["compliance"] - row["compliance"]) <= 1e-12 * abs(row["compliance"])
        return
    pytest.skip("no state kept the baseline floor")


def test_trained_network_gains_over_the_plain_policy():
    """The shipped trained parameters (bench.py --form learned, trained on
    the labels in gpu_fem/data) against the method without them, on fresh
    seeds the network never saw (40x20x20 Bernoulli and filtered-sigmoid
    states): the learned policy skips the 100-iteration probe, so it needs
    fewer FGMRES iterations in total; the guard still judges every solve, so
    every accepted solve meets the tolerance and every state the plain
    policy accepts is accepted."""
    import os

    from gpu_fem.learned import (DEFAULT_MODEL, FloorCNN, compare, label_states, run_learned,
                                 sample_states)

    if not os.path.exists(DEFAULT_MODEL):
        pytest.skip("no trained parameters shipped")
    model = FloorCNN.load()
    states = sample_states(8, 77, [(40, 20, 20)])
    plain = label_states(states, workers=4)
    learned, _ = run_learned(model, states, workers=4)
    plain_iters = sum(len(r["probe_history"]) - 1 + sum(len(h) - 1 for _, h in r["solve_history"])
                      for r in plain)
    learned_iters = sum(r["total_iters"] for r in learned)
    c = compare(plain, learned)
    print("plain iters", plain_iters, "learned iters", learned_iters, c)
    assert learned_iters < plain_iters
    assert c["learned_accepted"] >= c["plain_accepted"]
    assert c["max_accepted_true_rel_residual"] <= 1e-6


def test_out_of_range_grid_goes_through_the_probe():
    """A state on a grid outside the shipped network's training range is not
    proposed for: it takes the plain probe path and its row equals the plain
    policy's (decisions and iterations)."""
    from gpu_fem.learned import FloorCNN, label_states, run_learned, sample_states

    model = FloorCNN.load()
    assert model.trained_grids
    dims = (24, 12, 12)
    assert not model.in_range(dims)
    states = sample_states(2, 31, [dims], kinds=("bernoulli",), p_range=(0.3, 0.5))
    plain = label_states(states, workers=1)
    learned, _ = run_learned(model, states, workers=1)
    for p, q in zip(plain, learned):
        assert q["gate"] == "probe_out_of_range"
        for k in ("trigger", "recommended_rho_min", "solve_iters", "solve_converged"):
            assert p[k] == q[k], k


def test_learned_form_at_512k():
    """The shipped network on fresh 160x80x40 states (512,000 elements, one of
    its training grids; seeds it never saw, Bernoulli and filtered-sigmoid):
    ungated and with the grid's calibrated confidence gate.  Every solve the
    learned policy accepts passes the recomputed-residual guard, it accepts
    every state the plain policy accepts, and ungated it spends fewer FGMRES
    iterations than the plain policy (no 100-iteration probe)."""
    import os

    from gpu_fem.learned import DEFAULT_MODEL, FloorCNN, compare, label_states, run_learned, \
        sample_states

    if not os.path.exists(DEFAULT_MODEL):
        pytest.skip("no trained parameters shipped")
    model = FloorCNN.load()
    dims = (160, 80, 40)
    assert model.in_range(dims)
    states = sample_states(4, 4242, [dims])
    plain = label_states(states, workers=2)
    plain_iters = sum(len(r["probe_history"]) - 1 + sum(len(h) - 1 for _, h in r["solve_history"])
                      for r in plain)
    for gate in (0.0, "auto"):
        learned, _ = run_learned(model, states, workers=2, confidence_threshold=gate)
        c = compare(plain, learned)
        iters = sum(r["total_iters"] for r in learned)
        print(f"512k gate {gate}: plain iters {plain_iters}, learned iters {iters}, {c}",
              flush=True)
        assert c["learned_accepted"] >= c["plain_accepted"]
        assert c["max_accepted_true_rel_residual"] <= 1e-6
        if gate == 0.0:
            assert iters < plain_iters


def test_evaluate_reports_agreement_against_the_plain_policy():
    """The fresh-state evaluation runs the plain policy and both learned
    forms on the same states and compares them: every row of the learned
    runs is counted, the per-grid comparison adds up to the overall one,
    and proposing the plain floor state by state reproduces agreement 1."""
    from gpu_fem import learned as L

    model = L.FloorCNN.load()
    states = L.sample_states(4, 31, [(40, 20, 20), (20, 10, 10)])
    res = L.evaluate(model, states, workers=2)
    assert res["states"] == 4 and res["plain_fp64"]["accepted"] >= 1
    for name in ("fp32_ungated", "fp32_gated"):
        c = res[name]
        assert c["states"] == 4
        assert sum(g["states"] for g in c["per_grid"].values()) == 4
        agree = sum(g["floor_agreement"] * g["states"] for g in c["per_grid"].values())
        assert abs(agree - c["floor_agreement"] * 4) < 1e-9
        # a grid outside the training range always takes the plain probe
        small = c["per_grid"].get("20x10x10")
        if small is not None:
            assert small["out_of_range_rate"] == 1.0 and small["floor_agreement"] == 1.0
        assert c["max_accepted_true_rel_residual"] <= 1e-6
    # one record per state, in state order
    assert [r["seed"] for r in res["rows"]] == [s["seed"] for s in states]
    for r in res["rows"]:
        if r["gated_path"] == "proposal":
            assert r["confidence"] >= r["gate"]


def test_probed_states_are_solved_as_the_plain_policy_solves_them():
    """With the plain policy's FP64 stack, a state that the confidence gate
    sends to the probe gives the plain policy's row bit for bit (trigger,
    floor, probe and solve histories), also when the session's first state
    is solved at a proposed floor: the session sets that instance up with
    the probe's moduli first, so its cached spectral estimate is the plain
    session's."""
    from gpu_fem import learned as L

    model = L.FloorCNN.load()
    states = L.sample_states(8, 43, [(40, 20, 20)])
    _, _, conf = L.predict_states(model, states)
    conf = np.asarray(conf, dtype=np.float64)
    order = np.argsort(-conf, kind="stable")  # the most confident state first
    states = [states[i] for i in order]
    conf = conf[order]
    tau = float(np.median(conf))
    assert conf[0] > tau
    plain = L.label_states(states, workers=2)
    learned, _ = L.run_learned(model, states, workers=2, confidence_threshold=tau)
    assert learned[0]["gate"] == "proposal"
    probed = [i for i, r in enumerate(learned) if r["gate"] == "probe"]
    assert len(probed) >= 3
    for i in probed:
        p, q = plain[i], learned[i]
        assert q["trigger"] == p["trigger"]
        assert q["recommended_rho_min"] == p["recommended_rho_min"]
        assert q["probe_history"] == p["probe_history"]
        assert q["solve_history"] == p["solve_history"]
