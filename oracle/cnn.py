"""Oracle: the learned floor predictor (NumPy).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

The reference has no learned predictor; this module states the network the
learned-floor form adds on top of the guarded policy, and the exact
arithmetic of its BF16 training so that the HIP kernels can be checked:

  input   density average-pooled to 16 x 8 x 8 (adaptive bins
          [floor(i n / m), ceil((i + 1) n / m)) per axis, as
          torch.nn.AdaptiveAvgPool3d) plus three grid features log2(dim) - 5
  conv1   3x3x3, zero padding 1, 1 -> 16 channels, ReLU
  conv2   3x3x3, zero padding 1, 16 -> 32 channels, ReLU
  pool    global average over the 1024 positions -> 32
  fc3     [32 pooled, 3 grid features] -> 32, ReLU
  fc4     32 -> 4: output 0 regresses the normalised log10 r50, outputs 1..3
          are the logits of the floor class (index into the floor ladder)
  loss    mean over the batch of w_reg (out0 - y_reg)^2 + cross-entropy

BF16 rule: every product of two tensors (the convolutions and the dense
layers, forward and backward) takes its operands rounded to bfloat16 (round
to nearest even from the FP32 value) and accumulates in FP32; biases, ReLU,
pooling, the loss and the optimizer stay FP32.  ``bf16=False`` turns every
rounding off (float64 throughout).

Parameter vector (float32, 15,588 values), in this order:
  W1 [16][27]        (tap t = (kd*3 + kh)*3 + kw)
  b1 [16]
  W2 [32][27][16]    (out channel, tap, in channel)
  b2 [32]
  W3 [32][35]
  b3 [32]
  W4 [4][32]
  b4 [4]

Pinning (``tests/test_oracle_cnn.py``):
* with ``bf16=False`` the forward pass, the loss and every parameter
  gradient equal torch (CPU, float64) conv3d / linear / relu / mean /
  cross_entropy with autograd, to 1e-12;
* ``adaptive_pool`` equals torch.nn.functional.adaptive_avg_pool3d on
  grids that divide unevenly;
* ``bf16_round`` equals torch's float32 -> bfloat16 conversion bit for bit
  (ties, subnormals, signs);
* ``adam_step`` equals torch.optim.Adam for several steps;
* central finite differences of the float64 loss match the gradient.
The BF16 rounding points themselves are this module's definition (nothing
in the reference fixes them); the HIP kernels are held to them.
"""

from __future__ import annotations

import numpy as np

POOL = (16, 8, 8)
NPOS = 16 * 8 * 8
C1, C2, NTAP, NFEAT, NHID, NOUT = 16, 32, 27, 3, 32, 4

_SIZES = [("W1", (C1, NTAP)), ("b1", (C1,)), ("W2", (C2, NTAP, C1)), ("b2", (C2,)),
          ("W3", (NHID, C2 + NFEAT)), ("b3", (NHID,)), ("W4", (NOUT, NHID)), ("b4", (NOUT,))]
OFFSETS = {}
_o = 0
for _name, _shape in _SIZES:
    OFFSETS[_name] = (_o, _shape)
    _o += int(np.prod(_shape))
N_PARAMS = _o  # 15588


def unpack(theta):
    theta = np.asarray(theta)
    return {k: theta[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in OFFSETS.items()}


def pack(p: dict) -> np.ndarray:
    out = np.zeros(N_PARAMS, dtype=np.float64)
    for k, (o, s) in OFFSETS.items():
        out[o:o + int(np.prod(s))] = np.asarray(p[k], dtype=np.float64).reshape(-1)
    return out


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights (bound sqrt(6 / fan_in)), zero biases, float32."""
    rng = np.random.default_rng(seed)
    p = {}
    for k, (o, s) in OFFSETS.items():
        if k.startswith("b"):
            p[k] = np.zeros(s)
        else:
            fan_in = int(np.prod(s[1:]))
            bound = np.sqrt(6.0 / fan_in)
            p[k] = rng.uniform(-bound, bound, size=s)
    return pack(p).astype(np.float32)


def bf16_round(x) -> np.ndarray:
    """Round to bfloat16 (nearest even) from the float32 value; returned as
    float64.  NaN/Inf are not expected here."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    b = f.view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return (b & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def _r(x, on):
    return bf16_round(x) if on else np.asarray(x, dtype=np.float64)


def adaptive_pool(rho, dims, out=POOL) -> np.ndarray:
    """Average of the density over adaptive bins (element order e = ix*ny*nz
    + iy*nz + iz)."""
    nx, ny, nz = (int(v) for v in dims)
    r = np.asarray(rho, dtype=np.float64).reshape(nx, ny, nz)
    res = np.zeros(out)
    for i in range(out[0]):
        x0, x1 = (i * nx) // out[0], -((-(i + 1) * nx) // out[0])
        for j in range(out[1]):
            y0, y1 = (j * ny) // out[1], -((-(j + 1) * ny) // out[1])
            for k in range(out[2]):
                z0, z1 = (k * nz) // out[2], -((-(k + 1) * nz) // out[2])
                res[i, j, k] = r[x0:x1, y0:y1, z0:z1].mean()
    return res


def grid_features(dims) -> np.ndarray:
    return np.log2(np.asarray(dims, dtype=np.float64)) - 5.0


def _patches(a):
    """a: [B, C, 16, 8, 8] -> [B, 27, C, 1024]: the zero-padded 3x3x3
    neighbourhood of every position, tap-major."""
    B, C = a.shape[:2]
    p = np.zeros((B, C, 18, 10, 10))
    p[:, :, 1:17, 1:9, 1:9] = a
    out = np.empty((B, NTAP, C, NPOS))
    for kd in range(3):
        for kh in range(3):
            for kw in range(3):
                t = (kd * 3 + kh) * 3 + kw
                out[:, t] = p[:, :, kd:kd + 16, kh:kh + 8, kw:kw + 8].reshape(B, C, NPOS)
    return out


def _unpatch(g):
    """Adjoint of _patches: [B, 27, C, 1024] -> [B, C, 16, 8, 8]."""
    B, _, C, _ = g.shape
    p = np.zeros((B, C, 18, 10, 10))
    for kd in range(3):
        for kh in range(3):
            for kw in range(3):
                t = (kd * 3 + kh) * 3 + kw
                p[:, :, kd:kd + 16, kh:kh + 8, kw:kw + 8] += g[:, t].reshape(B, C, 16, 8, 8)
    return p[:, :, 1:17, 1:9, 1:9]


def forward(theta, X, feat, bf16=True):
    """X [B, 16, 8, 8], feat [B, 3] -> (out [B, 4], cache)."""
    p = unpack(np.asarray(theta, dtype=np.float64))
    X = np.asarray(X, dtype=np.float64)
    B = X.shape[0]
    P1 = _patches(_r(X, bf16).reshape(B, 1, 16, 8, 8))            # [B,27,1,1024]
    Z1 = np.einsum("ot,btp->bop", _r(p["W1"], bf16), P1[:, :, 0]) + p["b1"][None, :, None]
    A1 = np.maximum(Z1, 0.0)                                      # [B,16,1024]
    P2 = _patches(_r(A1, bf16).reshape(B, C1, 16, 8, 8))          # [B,27,16,1024]
    Z2 = np.einsum("otc,btcp->bop", _r(p["W2"], bf16), P2) + p["b2"][None, :, None]
    A2 = np.maximum(Z2, 0.0)
    g = A2.mean(axis=2)                                           # [B,32]
    h = np.concatenate([g, np.asarray(feat, dtype=np.float64)], axis=1)  # [B,35]
    Z3 = _r(h, bf16) @ _r(p["W3"], bf16).T + p["b3"]
    A3 = np.maximum(Z3, 0.0)
    out = _r(A3, bf16) @ _r(p["W4"], bf16).T + p["b4"]
    cache = dict(p=p, P1=P1, Z1=Z1, A1=A1, P2=P2, Z2=Z2, h=h, Z3=Z3, A3=A3, B=B)
    return out, cache


def loss_and_grad(theta, X, feat, y_reg, y_cls, w_reg=1.0, bf16=True):
    """Mean loss and its gradient with respect to theta (float64 arrays)."""
    out, c = forward(theta, X, feat, bf16)
    p, B = c["p"], c["B"]
    y_reg = np.asarray(y_reg, dtype=np.float64)
    y_cls = np.asarray(y_cls, dtype=np.int64)
    r = out[:, 0] - y_reg
    lg = out[:, 1:]
    m = lg.max(axis=1, keepdims=True)
    lse = m[:, 0] + np.log(np.exp(lg - m).sum(axis=1))
    losses = w_reg * r * r + lse - lg[np.arange(B), y_cls]
    sm = np.exp(lg - lse[:, None])
    dout = np.zeros_like(out)
    dout[:, 0] = 2.0 * w_reg * r / B
    dout[:, 1:] = sm / B
    dout[np.arange(B), 1 + y_cls] -= 1.0 / B
    gr = {}
    dR = _r(dout, bf16)
    gr["W4"] = dR.T @ _r(c["A3"], bf16)
    gr["b4"] = dout.sum(axis=0)
    dZ3 = (dR @ _r(p["W4"], bf16)) * (c["Z3"] > 0)
    d3R = _r(dZ3, bf16)
    gr["W3"] = d3R.T @ _r(c["h"], bf16)
    gr["b3"] = dZ3.sum(axis=0)
    dg = (d3R @ _r(p["W3"], bf16))[:, :C2]
    dZ2 = (dg[:, :, None] / NPOS) * (c["Z2"] > 0)                 # [B,32,1024]
    d2R = _r(dZ2, bf16)
    gr["W2"] = np.einsum("bop,btcp->otc", d2R, c["P2"])
    gr["b2"] = dZ2.sum(axis=(0, 2))
    dP2 = np.einsum("otc,bop->btcp", _r(p["W2"], bf16), d2R)
    dA1 = _unpatch(dP2).reshape(B, C1, NPOS)
    dZ1 = dA1 * (c["Z1"] > 0)
    gr["W1"] = np.einsum("bop,btp->ot", _r(dZ1, bf16), c["P1"][:, :, 0])
    gr["b1"] = dZ1.sum(axis=(0, 2))
    return float(losses.mean()), pack(gr), out


def adam_step(theta, m, v, g, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba, Algorithm 1) in float64; returns new
    (theta, m, v).  ``step`` counts from 1."""
    g = np.asarray(g, dtype=np.float64)
    m = beta1 * np.asarray(m, dtype=np.float64) + (1.0 - beta1) * g
    v = beta2 * np.asarray(v, dtype=np.float64) + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return np.asarray(theta, dtype=np.float64) - lr * mh / (np.sqrt(vh) + eps), m, v
