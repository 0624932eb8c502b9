"""Matrix-free fine operator, node-block Jacobi and PCG over HIP kernels.

Drop-in counterparts of solver_v2.py of the reference:
  MatrixFreeKff      solver_v2.py:1121-1251
  NodeBlockJacobiMF  solver_v2.py:1268-1451
  _cupy_pcg          solver_v2.py:551-623
Device arrays are torch tensors on the ROCm device (the reference used CuPy
arrays); all arithmetic runs in libgfhip.so.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

from . import _lib
from .pub_simp_solver import grid_from_edof


def _host(a, dtype=None):
    """numpy copy of a torch tensor / numpy array / sequence."""
    try:
        import torch

        if isinstance(a, torch.Tensor):
            a = a.detach().cpu().numpy()
    except ImportError:
        pass
    a = np.asarray(a)
    return a.astype(dtype) if dtype is not None else a


def _as_dev(t, dtype):
    import torch

    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(np.asarray(t))
    if t.device.type != "cuda":
        t = t.to("cuda")
    if t.dtype != dtype:
        t = t.to(dtype)
    return t.contiguous()


class _Ctx:
    """Owns one gf_ctx (structured grid + free map + Ke on the device)."""

    def __init__(self, grid, KE, free, ndof):
        _lib.require_gpu()
        lib = _lib.load()
        self.grid = tuple(int(v) for v in grid)
        self.KE = np.ascontiguousarray(_host(KE, np.float64).reshape(24, 24))
        self.free = np.ascontiguousarray(_host(free, np.int32).ravel())
        self.ndof = int(ndof)
        h = C.c_void_p()
        _lib.check(lib.gf_ctx_create(C.byref(h), *self.grid, self.KE.ctypes.data,
                                     self.free.ctypes.data, len(self.free), self.ndof),
                   "gf_ctx_create")
        self.h = h
        self._lib = lib

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self._lib.gf_ctx_destroy(self.h)
                self.h = None
        except Exception:
            pass


def _tune_ctx(ctx: "_Ctx"):
    """Launch shape of a context's fine operator: GFHIP_FINE_SHAPE when set
    (profiling runs, where the tracer's per-dispatch cost can reorder the
    shapes' timings), else the autotuner's pick.  Returns (shape, timings
    in microseconds per shape or None)."""
    forced = int(os.environ.get("GFHIP_FINE_SHAPE", "0") or 0)
    if forced:
        if forced not in {code for _, code in LAUNCH_SHAPES}:
            raise ValueError("GFHIP_FINE_SHAPE must be one of "
                             f"{[c for _, c in LAUNCH_SHAPES]}; got {forced}")
        _lib.call("gf_ctx_set_shape", ctx.h, forced)
        return forced, None
    us = (C.c_float * len(LAUNCH_SHAPES))()
    shape = C.c_int32()
    _lib.call("gf_ctx_autotune", ctx.h, C.cast(us, C.c_void_p), C.byref(shape),
              _lib.stream_handle())
    return int(shape.value), {name: float(us[i]) for i, (name, _) in enumerate(LAUNCH_SHAPES)}


class FineOperator:
    """A_op = K_ff(E) bound to moduli E; a plain callable like the reference's
    ``lambda x: mf_op.matvec(x, E_e)``, recognisable by the fast solve path."""

    def __init__(self, mf_op, E):
        self.mf_op = mf_op
        self.E = E

    def __call__(self, x):
        return self.mf_op.matvec(x, self.E)


# fine-operator launch shapes in the autotuner's timing order: (name, code);
# the mfma shapes run the FP64 element-row contraction on the matrix cores
# (FP32: the same VALU tile); the box shapes stage a 3D node box per block
# and give each node one lane (box2x8x8s2: two)
LAUNCH_SHAPES = (("node", 1), ("wave4", 4), ("tile8x8", 8), ("tile4x16", 16),
                 ("mfma8x8", 72), ("mfma4x16", 80), ("box4x8x8", 48), ("box2x8x16", 50),
                 ("box2x8x8s2", 44))


class MatrixFreeKff:
    """y_free = K_ff(E) u_free without assembling K (solver_v2.py:1121-1251)."""

    def __init__(self, edof_gpu, KE_unit_gpu, free_gpu, n_free: int, ndof: int, *, grid=None,
                 autotune: bool = True):
        self._edof_gpu = edof_gpu
        self._KE_unit = KE_unit_gpu
        self._free_gpu = free_gpu
        self._n_free = int(n_free)
        self._ndof = int(ndof)
        self._n_elem, self._ke_size = int(edof_gpu.shape[0]), int(edof_gpu.shape[1])
        if grid is None:
            grid = grid_from_edof(_host(edof_gpu))
        self.grid = tuple(int(g) for g in grid)
        if self.grid[0] * self.grid[1] * self.grid[2] != self._n_elem:
            raise ValueError("grid does not match the element count")
        free = _host(free_gpu, np.int32)
        if len(free) != self._n_free:
            raise ValueError("len(free_gpu) != n_free")
        self._ctx = _Ctx(self.grid, KE_unit_gpu, free, ndof)
        self._KE_diag = np.diag(self._ctx.KE).copy()
        self.launch_shape = None
        self.launch_timings_us = None
        if autotune or os.environ.get("GFHIP_FINE_SHAPE"):
            self.launch_shape, self.launch_timings_us = _tune_ctx(self._ctx)

    def autotune(self):
        """Pick the fastest fine-operator launch shape for this grid (results
        are bitwise identical for every shape)."""
        self.launch_shape, self.launch_timings_us = _tune_ctx(self._ctx)
        return self.launch_shape

    def set_launch_shape(self, shape: int):
        """Force a launch shape (0 default, 1 node, 4 wave4, 8/16 LDS tiles,
        72/80 the LDS tiles with the FP64 contraction on the matrix cores,
        48/50 the boxes of 4x8x8 and 2x8x16 nodes with a lane per node, 44
        the 2x8x8 box with two lanes per node)."""
        _lib.call("gf_ctx_set_shape", self._ctx.h, int(shape))
        self.launch_shape = int(shape)

    @property
    def ctx_handle(self):
        return self._ctx.h

    def operator(self, E_e) -> FineOperator:
        return FineOperator(self, E_e)

    def matvec(self, u_free, E_e):
        import torch

        dt = u_free.dtype
        if dt not in (torch.float64, torch.float32):
            raise TypeError("u_free must be float64 or float32")
        u = _as_dev(u_free, dt)
        E = _as_dev(E_e, dt)
        if u.numel() != self._n_free or E.numel() != self._n_elem:
            raise ValueError("matvec: size mismatch")
        y = torch.empty_like(u)
        fn = "gf_matvec_f64" if dt == torch.float64 else "gf_matvec_f32"
        _lib.call(fn, self._ctx.h, _lib.ptr(E), _lib.ptr(u), _lib.ptr(y), _lib.stream_handle())
        return y

    def residual(self, b, u_free, E_e):
        """b - K_ff(E) u in one pass (FP64)."""
        import torch

        b = _as_dev(b, torch.float64)
        u = _as_dev(u_free, torch.float64)
        E = _as_dev(E_e, torch.float64)
        y = torch.empty_like(u)
        _lib.call("gf_residual_f64", self._ctx.h, _lib.ptr(E), _lib.ptr(b), _lib.ptr(u),
                  _lib.ptr(y), _lib.stream_handle())
        return y

    def extract_diagonal(self, E_e):
        import torch

        E = _as_dev(E_e, torch.float64)
        d = torch.empty(self._n_free, dtype=torch.float64, device=E.device)
        _lib.call("gf_diag_f64", self._ctx.h, _lib.ptr(E), _lib.ptr(d), _lib.stream_handle())
        return d


class NodeBlockJacobiMF:
    """3x3 node-block Jacobi (solver_v2.py:1268-1451)."""

    def __init__(self, edof_gpu, KE_unit_gpu, free_gpu, n_free: int, ndof: int, *, grid=None):
        self._n_free = int(n_free)
        self._ndof = int(ndof)
        self._n_nodes = ndof // 3
        if grid is None:
            grid = grid_from_edof(_host(edof_gpu))
        self.grid = tuple(int(g) for g in grid)
        free = _host(free_gpu, np.int32)
        self._ctx = _Ctx(self.grid, KE_unit_gpu, free, ndof)
        mask = np.zeros(ndof, dtype=bool)
        mask[free] = True
        nf = mask.reshape(-1, 3)
        self._n_all_free = int(nf.all(axis=1).sum())
        self._updated = False

    def update(self, E_e, scalar_diag=None) -> None:
        import torch

        E = _as_dev(E_e, torch.float64)
        sd = None if scalar_diag is None else _as_dev(scalar_diag, torch.float64)
        _lib.call("gf_nodeblock_update", self._ctx.h, _lib.ptr(E), _lib.ptr(sd),
                  _lib.stream_handle())
        self._updated = True

    def apply(self, r):
        import torch

        if not self._updated:
            raise RuntimeError("NodeBlockJacobiMF.update() must be called before apply()")
        dt = r.dtype
        rr = _as_dev(r, torch.float64)
        z = torch.empty_like(rr)
        _lib.call("gf_nodeblock_apply", self._ctx.h, _lib.ptr(rr), _lib.ptr(z),
                  _lib.stream_handle())
        return z.to(dt)


# ---------------------------------------------------------------------------
# BLAS-1 helpers on libgfhip (deterministic reductions)
# ---------------------------------------------------------------------------
_scalar_tls = threading.local()


def _scalar_buf(device):
    """A 2-double device buffer for scalar results, one per thread and device
    (threads drive independent solves on their own streams)."""
    import torch

    cache = getattr(_scalar_tls, "bufs", None)
    if cache is None:
        cache = _scalar_tls.bufs = {}
    key = str(device)
    if key not in cache:
        cache[key] = torch.zeros(2, dtype=torch.float64, device=device)
    return cache[key]


def dot(x, y) -> float:
    """float(x . y) -- one host synchronisation (as cp.dot + float())."""
    s = _scalar_buf(x.device)
    _lib.call("gf_dot_f64", None, x.numel(), _lib.ptr(x), _lib.ptr(y), _lib.ptr(s),
              _lib.stream_handle())
    return float(s[0].item())


def nrm2(x) -> float:
    s = _scalar_buf(x.device)
    _lib.call("gf_nrm2_f64", None, x.numel(), _lib.ptr(x), _lib.ptr(s), _lib.stream_handle())
    return float(s[0].item())


def axpby(a: float, x, b: float, y):
    """y <- a x + b y in place."""
    _lib.call("gf_axpby_f64", y.numel(), float(a), _lib.ptr(x), float(b), _lib.ptr(y),
              _lib.stream_handle())
    return y


def _cupy_pcg(A, b, M_op, x0=None, tol=1e-5, maxiter=1000, history=None):
    """PCG with the up-front convergence check (solver_v2.py:551-623)."""
    import torch

    Aop = A if callable(A) else (lambda v, _A=A: _A @ v)
    b = _as_dev(b, torch.float64)
    x = x0.clone() if x0 is not None else torch.zeros_like(b)
    r = b.clone()
    axpby(-1.0, Aop(x), 1.0, r)
    bb = dot(b, b)
    tol_sq = tol ** 2 * bb
    rr0 = dot(r, r)
    if history is not None:
        history.append((rr0 / bb) ** 0.5 if bb > 0 else 0.0)
    if rr0 <= tol_sq:
        return x, 0, True
    z = M_op(r)
    p = z.clone()
    rz = dot(r, z)
    converged = False
    iters = 1
    for _ in range(maxiter - 1):
        Ap = Aop(p)
        alpha = rz / dot(p, Ap)
        axpby(alpha, p, 1.0, x)
        axpby(-alpha, Ap, 1.0, r)
        rr = dot(r, r)
        if history is not None:
            history.append((rr / bb) ** 0.5 if bb > 0 else 0.0)
        if rr <= tol_sq:
            converged = True
            break
        z = M_op(r)
        iters += 1
        rz_new = dot(r, z)
        beta = rz_new / rz
        p = axpby(1.0, z, beta, p)
        rz = rz_new
    return x, iters, converged
