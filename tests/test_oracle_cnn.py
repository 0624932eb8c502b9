"""Pins the learned-floor network oracle (oracle/cnn.py) against torch on the
CPU (float64 autograd, bfloat16 conversion, Adam) and finite differences."""

import numpy as np
import pytest

from oracle import cnn

torch = pytest.importorskip("torch")
F = torch.nn.functional


def _torch_model(theta):
    p = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True)
         for k, v in cnn.unpack(np.asarray(theta, dtype=np.float64)).items()}
    return p


def _torch_forward(p, X, feat):
    B = X.shape[0]
    W1 = p["W1"].reshape(16, 1, 3, 3, 3)
    # ours: W2[co][t][ci] -> torch [co][ci][kd][kh][kw]
    W2 = p["W2"].permute(0, 2, 1).reshape(32, 16, 3, 3, 3)
    a1 = F.relu(F.conv3d(X.reshape(B, 1, 16, 8, 8), W1, p["b1"], padding=1))
    a2 = F.relu(F.conv3d(a1, W2, p["b2"], padding=1))
    g = a2.mean(dim=(2, 3, 4))
    h = torch.cat([g, feat], dim=1)
    a3 = F.relu(h @ p["W3"].T + p["b3"])
    return a3 @ p["W4"].T + p["b4"]


def _batch(B, seed):
    rng = np.random.default_rng(seed)
    X = rng.random((B, 16, 8, 8))
    feat = rng.uniform(-1, 2, size=(B, 3))
    y_reg = rng.normal(size=B)
    y_cls = rng.integers(0, 3, size=B)
    return X, feat, y_reg, y_cls


def test_parameter_layout():
    assert cnn.N_PARAMS == 16 * 27 + 16 + 32 * 27 * 16 + 32 + 32 * 35 + 32 + 4 * 32 + 4 == 15588


def test_forward_loss_and_gradient_match_torch_autograd():
    theta = cnn.init_params(3).astype(np.float64)
    theta += 0.05 * np.random.default_rng(9).normal(size=theta.shape)  # non-zero biases
    X, feat, y_reg, y_cls = _batch(5, 1)
    w_reg = 0.7
    loss, grad, out = cnn.loss_and_grad(theta, X, feat, y_reg, y_cls, w_reg, bf16=False)
    p = _torch_model(theta)
    Xt, ft = torch.tensor(X), torch.tensor(feat)
    out_t = _torch_forward(p, Xt, ft)
    lt = (w_reg * (out_t[:, 0] - torch.tensor(y_reg)) ** 2).mean() + \
        F.cross_entropy(out_t[:, 1:], torch.tensor(y_cls))
    lt.backward()
    assert np.abs(out - out_t.detach().numpy()).max() <= 1e-12 * max(1.0, np.abs(out).max())
    assert abs(loss - lt.item()) <= 1e-12 * max(1.0, abs(loss))
    gt = cnn.pack({k: v.grad.numpy() for k, v in p.items()})
    scale = np.abs(gt).max()
    assert np.abs(grad - gt).max() <= 1e-11 * scale
    for k, (o, s) in cnn.OFFSETS.items():  # every block of parameters is exercised
        assert np.abs(gt[o:o + int(np.prod(s))]).max() > 0, k


def test_gradient_central_differences():
    theta = cnn.init_params(5).astype(np.float64)
    X, feat, y_reg, y_cls = _batch(3, 2)
    _, grad, _ = cnn.loss_and_grad(theta, X, feat, y_reg, y_cls, 1.0, bf16=False)
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(np.arange(o, o + int(np.prod(s))), 3, replace=False)
                          for o, s in cnn.OFFSETS.values()])
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (cnn.loss_and_grad(tp, X, feat, y_reg, y_cls, 1.0, bf16=False)[0]
              - cnn.loss_and_grad(tm, X, feat, y_reg, y_cls, 1.0, bf16=False)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1e-3, abs(grad[i])), (i, fd, grad[i])


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(7)
    vals = np.concatenate([
        rng.normal(size=4000).astype(np.float32),
        (rng.normal(size=1000) * 1e-39).astype(np.float32),            # subnormals
        np.array([0.0, -0.0, 1.0, -1.0, 1.00390625, 1.01171875, 3.0e38], np.float32),
        # exact ties: low 16 bits 0x8000 with even and odd kept bit
        np.array([0x3F808000, 0x3F818000, 0xBF808000, 0x00008000], np.uint32).view(np.float32),
    ])
    ours = cnn.bf16_round(vals)
    ref = torch.tensor(vals).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(ours.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("dims", [(40, 20, 20), (30, 15, 15), (33, 9, 17), (16, 8, 8),
                                  (160, 80, 40)])
def test_adaptive_pool_matches_torch(dims):
    rho = np.random.default_rng(sum(dims)).random(int(np.prod(dims)))
    ours = cnn.adaptive_pool(rho, dims)
    ref = F.adaptive_avg_pool3d(torch.tensor(rho.reshape(1, 1, *dims)), (16, 8, 8))
    assert np.abs(ours - ref.numpy()[0, 0]).max() <= 1e-14


def test_adam_matches_torch():
    rng = np.random.default_rng(4)
    theta = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(4)]
    pt = torch.tensor(theta.copy(), requires_grad=True)
    opt = torch.optim.Adam([pt], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    th, m, v = theta.copy(), np.zeros(50), np.zeros(50)
    for k, g in enumerate(grads, start=1):
        opt.zero_grad()
        pt.grad = torch.tensor(g)
        opt.step()
        th, m, v = cnn.adam_step(th, m, v, g, k, 3e-3)
    assert np.abs(th - pt.detach().numpy()).max() <= 1e-15 * 1e3


def test_bf16_rounding_changes_only_products():
    """bf16=True differs from float64 by the operand rounding only: on inputs
    that are already bf16-representable with weights that are too, the first
    layer's pre-activations agree exactly."""
    theta = cnn.bf16_round(cnn.init_params(1)).astype(np.float32)
    X = cnn.bf16_round(np.random.default_rng(2).random((2, 16, 8, 8)))
    _, c64 = cnn.forward(theta, X, np.zeros((2, 3)), bf16=False)
    _, c16 = cnn.forward(theta, X, np.zeros((2, 3)), bf16=True)
    assert np.abs(c64["Z1"] - c16["Z1"]).max() <= 1e-15
    assert np.abs(c64["Z2"] - c16["Z2"]).max() > 0  # A1 is rounded before conv2
