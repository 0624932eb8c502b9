"""Learned-floor form: a small 3D CNN proposes the floor (and r50) from the
density, so that the probe solve is skipped; the recomputed-residual guard
judges every proposal and a rejected one falls back up the plain ladder.

Network (kernels in csrc/gf_cnn.hip, C-ABI gf_cnn_*): density average-pooled
to 16x8x8 plus grid features log2(dim) - 5 -> conv 3^3 1->16 ReLU -> conv 3^3
16->32 ReLU -> global average pool -> dense 35->32 ReLU -> dense 32->4
(output 0: normalised log10 r50, outputs 1..3: logits over the floor ladder
FLOORS).  Trained on the GPU in BF16 (FP32 accumulation) with Adam, on
labels that the plain probe/ladder/guard policy (policy.py, the reference's
run_gmg_floor_detector_prospective.py:191-354 logic) produces for synthetic
Bernoulli and filtered-sigmoid density states.

The reference has no learned predictor; what is kept from it is the guard
(``||F - K(E) x|| / ||F|| <= tol`` is the only acceptance test) and the
ladder order.
"""

from __future__ import annotations

import ctypes as C
import math
import os
import time

import numpy as np

from . import _lib
from .density_field import filtered_sigmoid_density
from .policy import GuardedFloorPolicy, PolicySession, _sync
from .presets import cantilever

FLOORS = (1e-12, 1e-3, 1e-2)
POOL = (16, 8, 8)

# The shipped network (data/ beside this module): parameters, training grids
# and per-grid confidence gates in floor_cnn.npz, trained on the GPU by
#   python -m gpu_fem.learned train --labels gpu_fem/data/floor_labels_train.csv,\
#       gpu_fem/data/floor_labels_large.csv,gpu_fem/data/floor_labels_1m.csv \
#       --epochs 60 --out gpu_fem/data/floor_cnn.npz
# from labels the plain policy wrote (python -m gpu_fem.learned label ...):
# 40x20x20, 64x32x16, 80x40x20 (floor_labels_train.csv), 100x50x25,
# 128x64x32, 160x80x40 (floor_labels_large.csv), 160x80x40 and 200x100x50
# (floor_labels_1m.csv).
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_MODEL = os.path.join(DATA_DIR, "floor_cnn.npz")
DEFAULT_LABELS = os.path.join(DATA_DIR, "floor_labels_train.csv")
LABEL_FILES = [os.path.join(DATA_DIR, n) for n in ("floor_labels_train.csv",
                                                   "floor_labels_large.csv",
                                                   "floor_labels_1m.csv")]


R50_LOG10_RANGE = (-12.0, 0.5)


def r50_target(r50) -> np.ndarray:
    """Normalised regression target (log10 r50 + 5) / 3, with log10 r50
    clipped to R50_LOG10_RANGE: a probe that breaks down records r50 = 0 and
    one that diverges records inf, and neither may dominate the loss."""
    r = np.asarray(r50, dtype=np.float64)
    lg = np.log10(np.clip(np.nan_to_num(r, nan=10.0, posinf=10.0), 1e-300, None))
    return (np.clip(lg, *R50_LOG10_RANGE) + 5.0) / 3.0


def r50_from_target(t) -> np.ndarray:
    return 10.0 ** (3.0 * np.asarray(t, dtype=np.float64) - 5.0)


def grid_features(dims) -> np.ndarray:
    return (np.log2(np.asarray(dims, dtype=np.float64)) - 5.0).astype(np.float32)


def in_grid_range(trained_grids, dims) -> bool:
    """Whether every axis of ``dims`` lies within the per-axis range of the
    training grids (True when they are not recorded): the grid features
    log2(dim) - 5 of a grid outside that range are an extrapolation."""
    if not trained_grids:
        return True
    g = np.asarray(trained_grids)
    d = np.asarray(dims)
    return bool(np.all(d >= g.min(axis=0)) and np.all(d <= g.max(axis=0)))


def floor_class(rho_min: float, floors=FLOORS) -> int:
    """Index of rho_min in the ladder (nearest in log scale)."""
    return int(np.argmin([abs(math.log10(rho_min) - math.log10(f)) for f in floors]))


def learned_ladder(predicted_class: int, floors, attempt):
    """Guarded attempts starting at the proposed floor and continuing up the
    ladder while the guard rejects.  ``attempt(rho_min) -> bool`` (guard
    verdict).  Returns (attempted floors, accepted index or None,
    fallback = the proposal itself was rejected)."""
    if not 0 <= predicted_class < len(floors):
        raise ValueError("predicted class outside the floor ladder")
    attempted = []
    for k in range(predicted_class, len(floors)):
        attempted.append(floors[k])
        if attempt(floors[k]):
            return attempted, k, k != predicted_class
    return attempted, None, True


# --------------------------------------------------------------------------
# network
def _sizes():
    n = C.c_int32()
    w = C.c_int64()
    _lib.call("gf_cnn_sizes", C.byref(n), C.byref(w))
    return int(n.value), int(w.value)


def init_theta(seed: int) -> np.ndarray:
    """He-uniform weights (bound sqrt(6 / fan_in)) in parameter order, zero
    biases, float32."""
    rng = np.random.default_rng(seed)
    blocks = [((16, 27), 27), ((16,), 0), ((32, 27, 16), 432), ((32,), 0), ((32, 35), 35),
              ((32,), 0), ((4, 32), 32), ((4,), 0)]
    out = []
    for shape, fan_in in blocks:
        if fan_in:
            b = math.sqrt(6.0 / fan_in)
            out.append(rng.uniform(-b, b, size=shape).reshape(-1))
        else:
            out.append(np.zeros(int(np.prod(shape))))
    return np.concatenate(out).astype(np.float32)


class FloorCNN:
    """Parameters, Adam state and device scratch of the floor predictor."""

    def __init__(self, theta=None, seed: int = 0, trained_grids=None, gates=None):
        import torch

        _lib.require_gpu()
        # grids of the training states; a state on a grid outside their
        # per-axis range is not proposed for (the policy probes it instead)
        self.trained_grids = ([tuple(int(v) for v in g) for g in trained_grids]
                              if trained_grids is not None else None)
        # confidence gate per training grid: {dims: tau}
        self.gates = {tuple(int(v) for v in k): float(t) for k, t in (gates or {}).items()}
        self.n_params, self._wps = _sizes()
        th = init_theta(seed) if theta is None else np.asarray(theta, dtype=np.float32)
        if th.shape != (self.n_params,):
            raise ValueError(f"theta must have {self.n_params} entries")
        self.theta = torch.as_tensor(th.copy(), device="cuda")
        self.m = torch.zeros_like(self.theta)
        self.v = torch.zeros_like(self.theta)
        self.grad_buf = torch.zeros_like(self.theta)
        self.loss_buf = torch.zeros(1, dtype=torch.float32, device="cuda")
        self.step = 0
        self._work = None

    def _work_for(self, B: int):
        import torch

        need = B * self._wps
        if self._work is None or self._work.numel() < need:
            self._work = torch.empty(need, dtype=torch.float32, device="cuda")
        return self._work

    @staticmethod
    def _check(X, feat):
        if X.dtype.itemsize != 4 or feat.dtype.itemsize != 4:
            raise ValueError("X and feat must be float32")
        B = X.shape[0]
        if X.numel() != B * 1024 or feat.numel() != B * 3:
            raise ValueError("X must be [B, 16, 8, 8] and feat [B, 3]")
        return B

    def forward(self, X, feat):
        """Outputs [B, 4] for pooled densities X [B, 16, 8, 8] and grid
        features feat [B, 3] (float32 device tensors)."""
        import torch

        B = self._check(X, feat)
        out = torch.empty((B, 4), dtype=torch.float32, device="cuda")
        _lib.call("gf_cnn_forward", _lib.ptr(self.theta), _lib.ptr(X.contiguous()),
                  _lib.ptr(feat.contiguous()), B, _lib.ptr(out), _lib.ptr(self._work_for(B)),
                  _lib.stream_handle())
        return out

    def grad(self, X, feat, y_reg, y_cls, w_reg: float = 1.0):
        """(mean loss [1], gradient [n_params], outputs [B, 4]) -- device
        tensors; the gradient buffer is reused by the next call."""
        import torch

        B = self._check(X, feat)
        out = torch.empty((B, 4), dtype=torch.float32, device="cuda")
        _lib.call("gf_cnn_grad", _lib.ptr(self.theta), _lib.ptr(X.contiguous()),
                  _lib.ptr(feat.contiguous()), _lib.ptr(y_reg.contiguous()),
                  _lib.ptr(y_cls.contiguous()), B, float(w_reg), _lib.ptr(self.loss_buf),
                  _lib.ptr(self.grad_buf), _lib.ptr(out), _lib.ptr(self._work_for(B)),
                  _lib.stream_handle())
        return self.loss_buf, self.grad_buf, out

    def adam(self, grad, lr: float, beta1: float = 0.9, beta2: float = 0.999,
             eps: float = 1e-8):
        self.step += 1
        _lib.call("gf_cnn_adam", _lib.ptr(self.theta), _lib.ptr(self.m), _lib.ptr(self.v),
                  _lib.ptr(grad), self.n_params, float(lr), float(beta1), float(beta2),
                  float(eps), self.step, _lib.stream_handle())

    def fit(self, X, feat, y_reg, y_cls, *, epochs: int = 100, batch: int = 64,
            lr: float = 3e-3, w_reg: float = 1.0, seed: int = 0, lr_final: float = 3e-4):
        """Minibatch Adam over the whole set (device tensors); per-epoch mean
        loss (host floats).  The learning rate decays geometrically from lr
        to lr_final.  Deterministic for a given seed."""
        import torch

        n = X.shape[0]
        rng = np.random.default_rng(seed)
        hist = []
        for ep in range(epochs):
            rate = lr * (lr_final / lr) ** (ep / max(1, epochs - 1))
            perm = torch.as_tensor(rng.permutation(n), device="cuda")
            acc = torch.zeros(1, dtype=torch.float32, device="cuda")
            nb = 0
            for s in range(0, n, batch):
                idx = perm[s:s + batch]
                if idx.numel() < 2:
                    continue
                loss, g, _ = self.grad(X.index_select(0, idx), feat.index_select(0, idx),
                                       y_reg.index_select(0, idx), y_cls.index_select(0, idx),
                                       w_reg)
                acc += loss
                nb += 1
                self.adam(g, rate)
            hist.append(float(acc.item()) / max(1, nb))
        return hist

    def predict(self, X, feat):
        """(floor class [B] int, predicted r50 [B]) on the host."""
        cls, r50, _ = self.predict_proba(X, feat)
        return cls, r50

    def predict_proba(self, X, feat):
        """(floor class [B], predicted r50 [B], confidence [B]) on the host;
        the confidence is the softmax probability of the chosen class."""
        out = self.forward(X, feat).double().cpu().numpy()
        z = out[:, 1:] - out[:, 1:].max(axis=1, keepdims=True)
        prob = np.exp(z) / np.exp(z).sum(axis=1, keepdims=True)
        return out[:, 1:].argmax(axis=1), r50_from_target(out[:, 0]), prob.max(axis=1)

    def theta_host(self) -> np.ndarray:
        return self.theta.cpu().numpy().copy()

    def in_range(self, dims) -> bool:
        """Whether the network may propose for a state on ``dims``."""
        return in_grid_range(self.trained_grids, dims)

    def gate_for(self, dims) -> float:
        """Confidence gate calibrated for the training grid nearest to
        ``dims`` in element count (0 when none is recorded)."""
        if not self.gates:
            return 0.0
        n = float(np.prod(dims))
        best = min(self.gates, key=lambda g: abs(math.log(np.prod(g) / n)))
        return self.gates[best]

    def save(self, path: str):
        """.npz: parameters, training grids and confidence gates; .npy:
        parameters only."""
        if path.endswith(".npz"):
            gk = sorted(self.gates)
            np.savez(path, theta=self.theta_host(),
                     trained_grids=np.asarray(self.trained_grids or [], dtype=np.int64),
                     gate_grids=np.asarray(gk, dtype=np.int64).reshape(-1, 3),
                     gate_tau=np.asarray([self.gates[k] for k in gk], dtype=np.float64))
        else:
            np.save(path, self.theta_host())

    @classmethod
    def load(cls, path: str = None):
        """Parameters saved by ``save`` (default: the shipped trained set).
        A bare .npy carries the parameters only (no training grids, no
        gates)."""
        path = path or DEFAULT_MODEL
        if path.endswith(".npz"):
            z = np.load(path, allow_pickle=False)
            grids = [tuple(int(v) for v in g) for g in z["trained_grids"]]
            gates = {tuple(int(v) for v in g): float(t)
                     for g, t in zip(z["gate_grids"], z["gate_tau"])} if "gate_tau" in z else {}
            return cls(theta=z["theta"], trained_grids=grids or None, gates=gates)
        return cls(theta=np.load(path, allow_pickle=False))


def pool_density(rho, dims):
    """Device float32 [16, 8, 8] pool of a density (host array or device
    float64 tensor, element order)."""
    import torch

    r = torch.as_tensor(np.asarray(rho, dtype=np.float64) if not torch.is_tensor(rho) else rho,
                        dtype=torch.float64, device="cuda").contiguous()
    nx, ny, nz = (int(v) for v in dims)
    if r.numel() != nx * ny * nz:
        raise ValueError("density size does not match the grid")
    X = torch.empty(POOL, dtype=torch.float32, device="cuda")
    _lib.call("gf_cnn_pool_f64", _lib.ptr(r), nx, ny, nz, _lib.ptr(X), _lib.stream_handle())
    return X


# --------------------------------------------------------------------------
# synthetic states
def sample_states(n: int, seed: int, grids, *, kinds=("bernoulli", "sigmoid"),
                  p_range=(0.1, 0.5), volfrac_range=(0.2, 0.5)):
    """Seeded state descriptors: grid drawn from ``grids``, density kind
    cycling over ``kinds``; Bernoulli solid probability or sigmoid volume
    fraction uniform in its range; density seed = 100000 * seed + index."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        dims = tuple(int(v) for v in grids[int(rng.integers(len(grids)))])
        kind = kinds[i % len(kinds)]
        st = {"dims": dims, "kind": kind, "seed": 100000 * seed + i}
        if kind == "bernoulli":
            st["probability"] = float(rng.uniform(*p_range))
        else:
            st["volfrac"] = float(rng.uniform(*volfrac_range))
        out.append(st)
    return out


def density_of(state) -> np.ndarray:
    """Element densities in [0, 1] of a state descriptor (flat)."""
    dims = state["dims"]
    if state["kind"] == "bernoulli":
        n_e = dims[0] * dims[1] * dims[2]
        return (np.random.default_rng(state["seed"]).random(n_e)
                < state["probability"]).astype(np.float64)
    return filtered_sigmoid_density(state["seed"], dims, volfrac=state["volfrac"])


def _policy_state(state, rho):
    if state["kind"] == "bernoulli":
        return {"solid": rho > 0.5, "seed": state["seed"],
                "probability": state.get("probability")}
    return {"rho_field": rho, "seed": state["seed"], "probability": None}


def _groups(states):
    order = {}
    for i, s in enumerate(states):
        order.setdefault(tuple(s["dims"]), []).append(i)
    return order


POLICY_DEFAULTS = dict(penal=3.0, raised_rho_mins=FLOORS[1:], baseline_rho_min=FLOORS[0],
                       probe_iters=100, probe_r50=50, restart=300, maxiter=300, tol=1e-6)


def _run_chunked(session, sts, chunk, progress, tag):
    """session.run over consecutive chunks (one session, so the cached
    lambda and the concurrency carry over).  ``progress`` gets a line after
    every chunk, and after any state that finishes 30 s or more after the
    previous line (long states on large grids)."""
    import threading
    import time

    out = []
    lock = threading.Lock()
    last = [time.perf_counter()]

    def on_done(_i):
        if progress is None:
            return
        with lock:
            now = time.perf_counter()
            if now - last[0] >= 30.0:
                last[0] = now
                progress(f"{tag}: a state finished ({len(out)}/{len(sts)} before this chunk)")

    for k in range(0, len(sts), chunk):
        out.extend(session.run(sts[k:k + chunk], on_done=on_done))
        if progress is not None:
            with lock:
                last[0] = time.perf_counter()
            progress(f"{tag}: {len(out)}/{len(sts)} states")
    return out


def label_states(states, *, workers: int = 12, load: str = "edge", densities=None,
                 progress=None, chunk: int = 48, **kw):
    """Plain guarded policy (probe -> classify -> ladder -> guard) on every
    state; one session per grid size.  Returns the policy rows in state
    order, each with ``dims``, ``kind`` and the pooled input ``X`` (host
    float32 [16, 8, 8]) added.  ``progress(msg)`` (optional) is called after
    every ``chunk`` states."""
    import torch

    pk = dict(POLICY_DEFAULTS)
    pk.update(kw)
    rows = [None] * len(states)
    for dims, idx in _groups(states).items():
        session = PolicySession(cantilever(*dims, load=load), workers=min(workers, len(idx)),
                                **pk)
        dens = [densities[i] if densities is not None else density_of(states[i]) for i in idx]
        got = _run_chunked(session, [_policy_state(states[i], d) for i, d in zip(idx, dens)],
                           chunk, progress, "plain policy " + "x".join(map(str, dims)))
        for i, d, r in zip(idx, dens, got):
            r.update(dims=dims, kind=states[i]["kind"],
                     X=pool_density(d, dims).cpu().numpy())
            r.pop("x", None)
            rows[i] = r
        del session
        torch.cuda.empty_cache()
    return rows


LABEL_FIELDS = ["index", "dims", "kind", "seed", "probability", "volfrac", "trigger",
                "probe_r50", "probe_r100", "recommended_rho_min", "n_floor_attempts",
                "solve_converged", "solve_iters", "recorded_policy_time_s"]


def write_labels(path: str, states, rows) -> None:
    """The plain policy's labels of ``states`` as CSV: the state descriptor
    (from which the density is regenerated) and the policy outcome."""
    import csv

    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=LABEL_FIELDS)
        w.writeheader()
        for i, (s, r) in enumerate(zip(states, rows)):
            w.writerow({
                "index": i, "dims": "x".join(str(d) for d in s["dims"]), "kind": s["kind"],
                "seed": s["seed"], "probability": s.get("probability", ""),
                "volfrac": s.get("volfrac", ""), "trigger": r["trigger"],
                "probe_r50": repr(float(r["probe_r50"])),
                "probe_r100": repr(float(r["probe_r100"])),
                "recommended_rho_min": repr(float(r["recommended_rho_min"])),
                "n_floor_attempts": r["n_floor_attempts"],
                "solve_converged": r["solve_converged"], "solve_iters": r["solve_iters"],
                "recorded_policy_time_s": repr(float(r["recorded_policy_time_s"])),
            })


def read_labels(path: str):
    """(states, rows) from a label CSV written by ``write_labels``; each row
    carries ``dims``, ``kind``, ``probe_r50``, ``recommended_rho_min`` and the
    policy columns (no pooled input: ``attach_pooled`` adds it)."""
    import csv

    states, rows = [], []
    with open(path) as f:
        for r in csv.DictReader(f):
            dims = tuple(int(v) for v in r["dims"].split("x"))
            st = {"dims": dims, "kind": r["kind"], "seed": int(r["seed"])}
            if r["kind"] == "bernoulli":
                st["probability"] = float(r["probability"])
            else:
                st["volfrac"] = float(r["volfrac"])
            states.append(st)
            rows.append({"dims": dims, "kind": r["kind"], "trigger": r["trigger"],
                         "probe_r50": float(r["probe_r50"]),
                         "probe_r100": float(r["probe_r100"]),
                         "recommended_rho_min": float(r["recommended_rho_min"]),
                         "n_floor_attempts": int(r["n_floor_attempts"]),
                         "solve_converged": int(r["solve_converged"]),
                         "solve_iters": int(r["solve_iters"]),
                         "recorded_policy_time_s": float(r["recorded_policy_time_s"])})
    return states, rows


def attach_pooled(states, rows) -> None:
    """Adds the pooled network input ``X`` (host float32 [16, 8, 8]) to each
    row, regenerating the density from the state descriptor."""
    for s, r in zip(states, rows):
        r["X"] = pool_density(density_of(s), s["dims"]).cpu().numpy()


def training_tensors(rows):
    """(X, feat, y_reg, y_cls) device tensors from labelled rows."""
    import torch

    X = np.stack([r["X"] for r in rows]).astype(np.float32)
    feat = np.stack([grid_features(r["dims"]) for r in rows])
    y_reg = r50_target([r["probe_r50"] for r in rows]).astype(np.float32)
    y_cls = np.array([floor_class(r["recommended_rho_min"]) for r in rows], dtype=np.int32)
    dev = lambda a: torch.as_tensor(a, device="cuda")  # noqa: E731
    return dev(X), dev(feat), dev(y_reg), dev(y_cls)


# --------------------------------------------------------------------------
# the learned policy
class LearnedFloorPolicy(GuardedFloorPolicy):
    """Solves a state at the network's proposed floor; the recomputed
    residual guard accepts it or the ladder continues upward (the plain
    ladder's raised floors).  With ``confidence_threshold`` > 0, a proposal
    whose confidence is below it is not used: that state goes through the
    plain probe -> classify -> ladder -> guard policy instead.  Same
    components and session semantics as GuardedFloorPolicy."""

    def __init__(self, preset_or_spec, *, floors=FLOORS, confidence_threshold: float = 0.0, **kw):
        kw.setdefault("baseline_rho_min", floors[0])
        kw.setdefault("raised_rho_mins", tuple(floors[1:]))
        super().__init__(preset_or_spec, **kw)
        self.floors = tuple(float(f) for f in floors)
        self.confidence_threshold = float(confidence_threshold)

    # a session's first state may be solved at its proposed floor first; the
    # session sets this instance up with the probe's moduli before it, so the
    # cached spectral estimate is the plain session's and a state sent to the
    # probe is solved exactly as the plain policy solves it (same stack)
    primes_first_state = True

    def _gated(self, confidence, out_of_range=False) -> bool:
        return bool(out_of_range) or float(confidence) < self.confidence_threshold

    def run_state(self, solid=None, *, rho_field=None, seed=None, probability=None,
                  keep_x=False, predicted_class=0, predicted_r50=float("nan"),
                  predicted_confidence=1.0, out_of_range=False):
        from .solver_v2 import dot

        if self._gated(predicted_confidence, out_of_range):
            row = super().run_state(solid, rho_field=rho_field, seed=seed,
                                    probability=probability, keep_x=keep_x)
            row.update(predicted_class=int(predicted_class),
                       predicted_rho_min=self.floors[int(predicted_class)],
                       predicted_r50=float(predicted_r50),
                       predicted_confidence=float(predicted_confidence),
                       gate="probe_out_of_range" if out_of_range else "probe",
                       guard_fallback=int(row["post_solve_fallback_triggered"]),
                       total_iters=(len(row["probe_history"]) - 1
                                    + sum(len(h) - 1 for _, h in row["solve_history"])))
            return row

        rec = {"x": None, "final": float("nan"), "iters": 0, "reported": False, "sel": None,
               "hist": [], "total_iters": 0, "time": 0.0}

        def attempt(rho_min):
            h: list = []
            _, x, it, rep, ok, fin, m = self.solve_at_floor(
                rho_min, maxiter=self.maxiter, history=h, solid=solid, rho_field=rho_field)
            rec.update(x=x, final=fin, iters=it, reported=rep, sel=m)
            rec["hist"].append((rho_min, h))
            rec["total_iters"] += len(h) - 1
            rec["time"] += m["setup_time_s"] + m["solve_time_s"]
            return ok

        attempted, acc, fallback = learned_ladder(int(predicted_class), self.floors, attempt)
        compliance = dot(self.F, rec["x"]) if rec["x"] is not None else float("nan")
        row = {
            "preset": self.preset, "seed": seed, "solid_probability": probability,
            "predicted_class": int(predicted_class),
            "predicted_rho_min": self.floors[int(predicted_class)],
            "predicted_r50": float(predicted_r50),
            "predicted_confidence": float(predicted_confidence), "gate": "proposal",
            "recommended_rho_min": attempted[-1],
            "attempted_rho_mins": ";".join(f"{v:.0e}" for v in attempted),
            "n_floor_attempts": len(attempted), "guard_fallback": int(fallback),
            "solve_converged": int(acc is not None),
            "solver_reported_converged": int(rec["reported"]),
            "solve_iters": rec["iters"], "total_iters": rec["total_iters"],
            "solve_final_rel_residual": rec["final"], "compliance": compliance,
            "recorded_policy_time_s": rec["time"], "penal": self.penal, "tol": self.tol,
            "restart": self.restart, "maxiter": self.maxiter,
            "fine_smoother": self.stack.get("fine_smoother"),
            "solve_history": rec["hist"],
        }
        if keep_x:
            row["x"] = rec["x"]
        return row


def confidence_for_accuracy(confidence, correct, target: float = 0.99) -> float:
    """Smallest confidence threshold tau such that the predictions with
    confidence >= tau are correct at least at rate ``target`` (measured on the
    given labelled set); 1.0 if no threshold reaches it."""
    conf = np.asarray(confidence, dtype=np.float64)
    ok = np.asarray(correct, dtype=bool)
    order = np.argsort(-conf)
    hits = np.cumsum(ok[order])
    rate = hits / np.arange(1, len(order) + 1)
    good = np.nonzero(rate >= target)[0]
    if len(good) == 0:
        return 1.0
    return float(conf[order][good[-1]])


def predict_states(model: FloorCNN, states, densities=None):
    """Pools every state's density on the device and runs one batched
    forward; (classes, r50s, confidences)."""
    import torch

    Xs, feats = [], []
    for i, s in enumerate(states):
        d = densities[i] if densities is not None else density_of(s)
        Xs.append(pool_density(d, s["dims"]))
        feats.append(torch.as_tensor(grid_features(s["dims"]), device="cuda"))
    return model.predict_proba(torch.stack(Xs), torch.stack(feats))


def run_learned(model: FloorCNN, states, *, workers: int = 12, load: str = "edge",
                densities=None, progress=None, chunk: int = 48, **kw):
    """Learned policy on every state (one session per grid size, concurrent
    instances).  Returns (rows in state order, wall seconds including the
    pooling and network inference).  ``progress(msg)`` (optional) is called
    after every ``chunk`` states."""
    import torch

    pk = dict(POLICY_DEFAULTS)
    pk.update(kw)
    dens = densities if densities is not None else [density_of(s) for s in states]
    _sync()
    t0 = time.perf_counter()
    cls, r50, conf = predict_states(model, states, dens)
    rows = [None] * len(states)
    for dims, idx in _groups(states).items():
        gk = dict(pk)
        if gk.get("confidence_threshold") == "auto":
            gk["confidence_threshold"] = model.gate_for(dims)
        session = PolicySession(cantilever(*dims, load=load), workers=min(workers, len(idx)),
                                runner_cls=LearnedFloorPolicy, **gk)
        sts = []
        for i in idx:
            st = _policy_state(states[i], dens[i])
            st["extra"] = {"predicted_class": int(cls[i]), "predicted_r50": float(r50[i]),
                           "predicted_confidence": float(conf[i]),
                           "out_of_range": not model.in_range(dims)}
            sts.append(st)
        got = _run_chunked(session, sts, chunk, progress,
                           "learned policy " + "x".join(map(str, dims)))
        for i, r in zip(idx, got):
            r.update(dims=dims, kind=states[i]["kind"])
            rows[i] = r
        del session
    torch.cuda.synchronize()
    return rows, time.perf_counter() - t0


def compare(plain_rows, learned_rows) -> dict:
    """Agreement of the accepted floor with the plain ladder, fallback rate,
    acceptance and worst accepted true residual."""
    n = len(plain_rows)
    agree = sum(abs(math.log10(p["recommended_rho_min"]) - math.log10(q["recommended_rho_min"]))
                < 1e-9 for p, q in zip(plain_rows, learned_rows))
    acc = [q for q in learned_rows if q["solve_converged"]]
    return {
        "states": n,
        "floor_agreement": agree / max(1, n),
        "proposal_agreement": sum(floor_class(p["recommended_rho_min"]) == q["predicted_class"]
                                  for p, q in zip(plain_rows, learned_rows)) / max(1, n),
        "fallback_rate": sum(q["guard_fallback"] for q in learned_rows) / max(1, n),
        "probe_rate": sum(q.get("gate", "").startswith("probe") for q in learned_rows)
        / max(1, n),
        "out_of_range_rate": sum(q.get("gate") == "probe_out_of_range" for q in learned_rows)
        / max(1, n),
        "learned_accepted": len(acc),
        "plain_accepted": sum(int(p["solve_converged"]) for p in plain_rows),
        "max_accepted_true_rel_residual": max([q["solve_final_rel_residual"] for q in acc],
                                              default=float("nan")),
        "r50_log10_rmse": float(np.sqrt(np.mean([
            (math.log10(max(q["predicted_r50"], 1e-300)) - math.log10(max(p["probe_r50"], 1e-300)))
            ** 2 for p, q in zip(plain_rows, learned_rows)]))),
    }


def calibrate_gates(model: "FloorCNN", rows, target: float = 0.99) -> dict:
    """Per-grid confidence gates: for each grid of ``rows`` (labelled states
    the network was not trained on), the smallest confidence at which its
    predictions are correct at rate ``target``."""
    X, feat, _, y_cls = training_tensors(rows)
    cls, _, conf = model.predict_proba(X, feat)
    ok = cls == y_cls.cpu().numpy()
    dims = [tuple(r["dims"]) for r in rows]
    gates = {}
    for g in sorted(set(dims)):
        sel = np.array([d == g for d in dims])
        gates[g] = confidence_for_accuracy(conf[sel], ok[sel], target)
    return gates


def evaluate(model, states, *, workers: int = 12, load: str = "edge", progress=None,
             chunk: int = 16, precision: str = "fp32", **kw) -> dict:
    """Fresh-state evaluation: the plain FP64 policy (the reference floor),
    then the learned policy with ``precision`` multigrid ("fp32" or "fp64")
    ungated and with the network's per-grid confidence gates, all on
    ``workers`` concurrent
    instances.  Returns the comparison of each learned run with the plain
    one, over all states and per grid, with accepted solves/s.  ``model``
    may be a dict of named networks evaluated on the same plain run; the
    result then holds one such comparison per name."""
    dens = [density_of(s) for s in states]

    def acc(rows):
        return sum(int(r["solve_converged"]) for r in rows)

    _sync()
    t0 = time.perf_counter()
    plain = label_states(states, workers=workers, load=load, densities=dens,
                         progress=progress, chunk=chunk, **kw)
    _sync()
    w_plain = time.perf_counter() - t0
    out = {"states": len(states),
           "grids": sorted({"x".join(map(str, s["dims"])) for s in states}),
           "plain_fp64": {"accepted": acc(plain), "wall_s": w_plain,
                          "value": acc(plain) / w_plain}}
    groups = _groups(states)
    if isinstance(model, dict):
        for mname, m in model.items():
            if progress is not None:
                progress(f"evaluating {mname}")
            out[mname] = _evaluate_learned(m, states, dens, plain, w_plain, groups, workers,
                                           load, progress, chunk, precision, kw)
        return out
    out.update(_evaluate_learned(model, states, dens, plain, w_plain, groups, workers, load,
                                 progress, chunk, precision, kw))
    return out


def _evaluate_learned(model, states, dens, plain, w_plain, groups, workers, load, progress,
                      chunk, precision, kw) -> dict:
    from .batch import precision_stack

    def acc(rows):
        return sum(int(r["solve_converged"]) for r in rows)

    out = {}
    runs = {}
    for name, gate in ((f"{precision}_ungated", 0.0), (f"{precision}_gated", "auto")):
        got, wall = run_learned(model, states, workers=workers, load=load, densities=dens,
                                progress=progress, chunk=chunk, confidence_threshold=gate,
                                **kw, **precision_stack(precision))
        runs[name] = got
        c = compare(plain, got)
        c.update(value=acc(got) / wall, wall_s=wall,
                 speedup_vs_plain_fp64_concurrent=(acc(got) / wall) / (acc(plain) / w_plain),
                 mean_total_iters=float(np.mean([r["total_iters"] for r in got])),
                 plain_mean_total_iters=float(np.mean(
                     [len(r["probe_history"]) - 1
                      + sum(len(h) - 1 for _, h in r["solve_history"]) for r in plain])))
        c["per_grid"] = {"x".join(map(str, d)): compare([plain[i] for i in idx],
                                                       [got[i] for i in idx])
                         for d, idx in groups.items()}
        out[name] = c
    # one record per state: the plain floor, the network's proposal and
    # confidence, the gate of its grid and what each learned run accepted
    # (the gated outcome for another gate follows from these: a state at or
    # above the gate takes the ungated run's result, one below it the plain
    # policy's)
    ug, ga = runs[f"{precision}_ungated"], runs[f"{precision}_gated"]
    out["rows"] = [{
        "dims": "x".join(map(str, s["dims"])), "kind": s["kind"], "seed": s["seed"],
        "plain_rho_min": p["recommended_rho_min"], "plain_trigger": p["trigger"],
        "plain_iters": len(p["probe_history"]) - 1
        + sum(len(h) - 1 for _, h in p["solve_history"]),
        "predicted_class": u["predicted_class"], "confidence": u["predicted_confidence"],
        "gate": model.gate_for(s["dims"]),
        "ungated_rho_min": u["recommended_rho_min"], "ungated_fallback": u["guard_fallback"],
        "ungated_iters": u["total_iters"],
        "gated_rho_min": g["recommended_rho_min"], "gated_path": g.get("gate", ""),
        "gated_iters": g["total_iters"],
    } for s, p, u, g in zip(states, plain, ug, ga)]
    return out


def main(argv=None):
    """Label states with the plain policy, or train the network.

    python -m gpu_fem.learned label --grids 160x80x40 --states 96 --seed 5 --out labels.csv
    python -m gpu_fem.learned train --labels a.csv,b.csv --out floor_cnn.npz
    python -m gpu_fem.learned evaluate --model floor_cnn.npz --grids 80x40x20 --states 96 \
        --seed 11 --out eval.json
    python -m gpu_fem.learned calibrate --model floor_cnn.npz --labels c.csv --target 0.995 \
        --out floor_cnn_cal.npz

    ``train`` holds out every seventh state of each grid to calibrate the
    per-grid confidence gates, trains on the rest (Adam, BF16 kernels) and
    saves the parameters with the training grids and the gates.  ``calibrate``
    sets the gates of a trained network afresh from labelled states it was
    not trained on (grids absent from those labels keep their gates)."""
    import argparse
    import sys

    import torch

    ap = argparse.ArgumentParser(prog="python -m gpu_fem.learned")
    sub = ap.add_subparsers(dest="cmd", required=True)
    la = sub.add_parser("label")
    la.add_argument("--grids", required=True)
    la.add_argument("--states", type=int, required=True)
    la.add_argument("--seed", type=int, required=True)
    la.add_argument("--workers", type=int, default=8)
    la.add_argument("--penal", type=float, default=3.0)
    la.add_argument("--out", required=True)
    ev = sub.add_parser("evaluate")
    ev.add_argument("--model", required=True)
    ev.add_argument("--grids", required=True)
    ev.add_argument("--states", type=int, required=True)
    ev.add_argument("--seed", type=int, required=True)
    ev.add_argument("--workers", type=int, default=12)
    ev.add_argument("--penal", type=float, default=3.0)
    ev.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ev.add_argument("--out", required=True)
    ca = sub.add_parser("calibrate")
    ca.add_argument("--model", required=True)
    ca.add_argument("--labels", required=True)
    ca.add_argument("--target", type=float, default=0.99)
    ca.add_argument("--out", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--labels", required=True)
    tr.add_argument("--epochs", type=int, default=60)
    tr.add_argument("--target", type=float, default=0.99)
    tr.add_argument("--out", required=True)
    a = ap.parse_args(argv)

    def say(msg):
        print(f"[learned {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)

    if a.cmd == "label":
        grids = [tuple(int(v) for v in g.split("x")) for g in a.grids.split(",") if g]
        states = sample_states(a.states, a.seed, grids)
        say(f"labelling {len(states)} states on {grids}")
        rows = label_states(states, workers=a.workers, progress=say, chunk=16, penal=a.penal)
        write_labels(a.out, states, rows)
        say(f"wrote {a.out}")
        return
    if a.cmd == "evaluate":
        import json

        grids = [tuple(int(v) for v in g.split("x")) for g in a.grids.split(",") if g]
        states = sample_states(a.states, a.seed, grids)
        paths = [m for m in a.model.split(",") if m]
        model = (FloorCNN.load(paths[0]) if len(paths) == 1
                 else {m: FloorCNN.load(m) for m in paths})
        say(f"evaluating {a.model} on {len(states)} fresh states on {grids}")
        res = evaluate(model, states, workers=a.workers, progress=say, penal=a.penal,
                       precision=a.precision)
        res.update(model=a.model, seed=a.seed, precision=a.precision)
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)
        say(f"wrote {a.out}")
        return
    states, rows = [], []
    for path in a.labels.split(","):
        s_, r_ = read_labels(path)
        states += s_
        rows += r_
    say(f"pooling {len(rows)} labelled states")
    attach_pooled(states, rows)
    if a.cmd == "calibrate":
        model = FloorCNN.load(a.model)
        gates = calibrate_gates(model, rows, a.target)
        say(f"gates at {a.target}: {gates} (before: {model.gates})")
        model.gates.update(gates)
        model.save(a.out)
        say(f"wrote {a.out}")
        return
    by_grid = {}
    for i, r in enumerate(rows):
        by_grid.setdefault(tuple(r["dims"]), []).append(i)
    hold = {i for idx in by_grid.values() for k, i in enumerate(idx) if k % 7 == 6}
    fit_rows = [r for i, r in enumerate(rows) if i not in hold]
    cal_rows = [r for i, r in enumerate(rows) if i in hold]
    X, feat, y_reg, y_cls = training_tensors(fit_rows)
    model = FloorCNN(seed=0, trained_grids=sorted(by_grid))
    say(f"training on {len(fit_rows)} states ({a.epochs} epochs)")
    hist = model.fit(X, feat, y_reg, y_cls, epochs=a.epochs, batch=64, lr=3e-3, seed=0)
    torch.cuda.synchronize()
    model.gates = calibrate_gates(model, cal_rows, a.target)
    cls, _, _ = model.predict_proba(X, feat)
    say(f"loss {hist[0]:.4f} -> {hist[-1]:.4f}; training accuracy "
        f"{float(np.mean(cls == y_cls.cpu().numpy())):.3f}; gates {model.gates}")
    model.save(a.out)
    say(f"wrote {a.out}")


if __name__ == "__main__":
    main()
