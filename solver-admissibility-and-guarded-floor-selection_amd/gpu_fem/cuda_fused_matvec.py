"""Fused hex8 matvec with FP32 and BF16 matrix-core forms.

Drop-in counterpart of cuda_fused_matvec.py:260-416 of the reference
(``FusedMatvec``).  The FP32 form runs the same owner-computes fine operator
as ``MatrixFreeKff`` in FP32 on the full DOF layout; the BF16 form rounds KE
and u to bf16 and contracts on gfx950 matrix cores
(v_mfma_f32_16x16x32_bf16) with FP32 accumulation.  Both are deterministic
(no atomics), unlike the reference's scatter-add.
"""

from __future__ import annotations

import numpy as np

from . import _lib
from .pub_simp_solver import grid_from_edof
from .solver_v2 import _Ctx, _as_dev, _host, _tune_ctx


class FusedMatvec:
    """y = K(E) u with a fused gather / Ke contraction / node sum."""

    KE_SIZE = 24
    BLOCK_X_FP32 = 256          # lanes per workgroup of the FP32 kernel
    BLOCK_X_BF16 = 256          # 4 waves x 16 elements per workgroup
    ELEMS_PER_BLOCK_BF16 = 64
    BLOCK_X = BLOCK_X_FP32

    def __init__(self, edof_gpu, KE_unit_gpu, ndof: int, *, grid=None):
        n_elem, ke_size = tuple(int(v) for v in edof_gpu.shape)
        if ke_size != self.KE_SIZE:
            raise ValueError(f"FusedMatvec currently supports ke_size=24 only, got {ke_size}")
        if grid is None:
            grid = grid_from_edof(_host(edof_gpu))
        self.grid = tuple(int(g) for g in grid)
        self._ndof = int(ndof)
        self._n_elem = n_elem
        # identity DOF map: the full layout of matvec_full
        self._ctx = _Ctx(self.grid, KE_unit_gpu, np.arange(self._ndof, dtype=np.int32),
                         self._ndof)
        # the FP32 form runs the fine operator's launch shape chosen for this
        # grid (bitwise the same result for every shape)
        self.launch_shape, self.launch_timings_us = _tune_ctx(self._ctx)
        self._bf16_available = True
        self._u_full = None
        self._y_full = None

    def _scratch(self, device):
        import torch

        if self._u_full is None:
            self._u_full = torch.zeros(self._ndof, dtype=torch.float32, device=device)
            self._y_full = torch.zeros(self._ndof, dtype=torch.float32, device=device)
        return self._u_full, self._y_full

    def _launch(self, dtype: str, u_full, y_full, E_f32):
        if dtype == "fp32":
            fn = "gf_matvec_full_f32"
        elif dtype == "bf16":
            fn = "gf_matvec_full_bf16"
        else:
            raise ValueError(f"Unknown dtype '{dtype}' — use 'fp32' or 'bf16'")
        _lib.call(fn, self._ctx.h, _lib.ptr(E_f32), _lib.ptr(u_full), _lib.ptr(y_full),
                  _lib.stream_handle())

    def matvec(self, u_free, E_e, free_gpu, *, dtype: str = "fp32"):
        """y_free = K_ff u_free in FP32 (cuda_fused_matvec.py:343-377)."""
        import torch

        u = _as_dev(u_free, torch.float32)
        E = _as_dev(E_e, torch.float32)
        free = _as_dev(free_gpu, torch.int64)
        u_full, y_full = self._scratch(u.device)
        u_full.zero_()
        u_full[free] = u
        self._launch(dtype, u_full, y_full, E)
        return y_full[free]

    def matvec_full(self, u_full_in, E_e, *, dtype: str = "fp32"):
        """y_full = K u_full over every DOF; u_full must be zero on fixed DOFs
        (cuda_fused_matvec.py:379-391)."""
        import torch

        u = _as_dev(u_full_in, torch.float32)
        E = _as_dev(E_e, torch.float32)
        if u.numel() != self._ndof or E.numel() != self._n_elem:
            raise ValueError("matvec_full: size mismatch")
        u_full, y_full = self._scratch(u.device)
        u_full.copy_(u)
        self._launch(dtype, u_full, y_full, E)
        return y_full.clone()
