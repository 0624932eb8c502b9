"""gf_residual_norm_f64_noalloc against gf_residual_f64 + gf_nrm2_f64 through
the C ABI: the same two norms bit for bit, on two calls in a row (the second
on a smaller grid), and a clear error above the scratch maximum."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _operator(dims, load="tip"):
    from gpu_fem import policy as gpol
    from gpu_fem.presets import cantilever

    comps = gpol._build_components(cantilever(*dims, load=load), n_levels=2)
    return comps[6], comps[5]


def _existing(lib, ctx, E, b, u):
    import torch
    from gpu_fem import _lib

    r = torch.empty_like(b)
    out = torch.empty(2, dtype=torch.float64, device="cuda")
    st = _lib.stream_handle()
    _lib.check(lib.gf_residual_f64(ctx, _lib.ptr(E), _lib.ptr(b), _lib.ptr(u), _lib.ptr(r), st))
    _lib.check(lib.gf_nrm2_f64(ctx, r.numel(), _lib.ptr(r), _lib.ptr(out[0:1]), st))
    _lib.check(lib.gf_nrm2_f64(ctx, b.numel(), _lib.ptr(b), _lib.ptr(out[1:2]), st))
    return out.cpu().numpy()


def _noalloc(lib, ctx, E, b, u):
    import torch
    from gpu_fem import _lib

    out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
    _lib.check(lib.gf_residual_norm_f64_noalloc(ctx, _lib.ptr(E), _lib.ptr(b), _lib.ptr(u),
                                                _lib.ptr(out), _lib.stream_handle()),
               "gf_residual_norm_f64_noalloc")
    return out.cpu().numpy()


def test_noalloc_guard_norms_equal_existing_entry_points():
    import torch
    from gpu_fem import _lib

    lib = _lib.load()
    rng = np.random.default_rng(3)
    for dims in [(80, 40, 20), (24, 12, 10)]:  # second call: smaller input
        mf, F = _operator(dims)
        n_e = dims[0] * dims[1] * dims[2]
        E = torch.as_tensor(1e-3 + rng.random(n_e), device="cuda")
        u = torch.as_tensor(rng.standard_normal(F.numel()), device="cuda")
        b = F + torch.as_tensor(1e-3 * rng.standard_normal(F.numel()), device="cuda")
        ref = _existing(lib, mf.ctx_handle, E, b, u)
        got = _noalloc(lib, mf.ctx_handle, E, b, u)
        print(f"{dims}: existing {ref.tolist()} no-alloc {got.tolist()}", flush=True)
        assert np.array_equal(ref, got), (dims, ref, got)
        # the guard's ratio against a plain float64 recomputation
        r = (b - mf.matvec(u, E)).cpu().numpy()
        assert abs(got[0] - np.linalg.norm(r)) <= 1e-12 * np.linalg.norm(r)
        assert abs(got[1] - np.linalg.norm(b.cpu().numpy())) <= 1e-13 * got[1]


def test_noalloc_rejects_contexts_above_the_scratch_maximum():
    import torch
    from gpu_fem import _lib
    from gpu_fem.pub_simp_solver import KE_UNIT_3D

    lib = _lib.load()
    nx, ny, nz = 120, 120, 100
    ndof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    free = np.arange(3 * (ny + 1) * (nz + 1), ndof, dtype=np.int32)
    assert len(free) > (1 << 22)
    h = C.c_void_p()
    ke = np.ascontiguousarray(KE_UNIT_3D, dtype=np.float64)
    _lib.check(lib.gf_ctx_create(C.byref(h), nx, ny, nz, ke.ctypes.data, free.ctypes.data,
                                 len(free), ndof), "gf_ctx_create")
    try:
        t = torch.zeros(4, dtype=torch.float64, device="cuda")
        rc = lib.gf_residual_norm_f64_noalloc(h, _lib.ptr(t), _lib.ptr(t), _lib.ptr(t),
                                              _lib.ptr(t), _lib.stream_handle())
        assert rc == -2
        assert b"exceed the scratch maximum" in lib.gf_last_error()
    finally:
        lib.gf_ctx_destroy(h)
