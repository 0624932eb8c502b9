"""Diagnostics through the C ABI against the oracle (run on the MI355X box).

``cycle_trace`` (gf_gmg_cycle_trace) field by field against the oracle's
restatement of multigrid_v4.py:1361-1728, for V and W cycles, both coarse
correction policies, root enrichment and both root-local modes: equal visit
structure, decisions and notes; every norm, ratio and scale to 1e-8 relative
(one cycle: the operators agree to 1e-13 and the nonlinear decisions are the
same).  Quantities that are zero in exact arithmetic (listed in the test) are
rounding noise in any implementation and are compared only for their size.
The traced output is the iterate apply() returns (bitwise) when there
is no inner Krylov wrapper.  ``estimate_kappa_eff`` against the oracle from
the same start vector to 1e-6 relative.
"""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BASE = dict(n_levels=4, fine_smoother="fp64", coarse_correction_policy="residual_line_search",
            root_local_correction_mode="node_block_inner_fgmres",
            root_local_correction_inner_steps=20, root_local_correction_inner_restart=20,
            inner_krylov_steps=0)


def _t(a):
    import torch

    return torch.as_tensor(np.asarray(a), device="cuda")


def _pair(dims, rho_min, **stack):
    from gpu_fem import policy as gpol
    from gpu_fem.presets import cantilever
    from oracle import fem, gmg as ogmg, policy as opol

    st = dict(BASE)
    st.update(stack)
    spec_o = fem.cantilever_spec(*dims, load="tip")
    bc, free, edof, F, mf = fem.build_components(spec_o)
    og = ogmg.GalerkinMatFreeGMG(mf_op=mf, free=free, nelx=dims[0], nely=dims[1], nelz=dims[2],
                                 KE_UNIT=fem.KE_UNIT_3D, **st)
    gg = gpol._build_components(cantilever(*dims, load="tip"), **st)[7]
    mask = opol.bernoulli_mask(0, dims, 0.35).ravel()
    E = opol.simp_modulus(np.where(mask, 1.0, rho_min), rho_min, 4.5)
    og.setup(E)
    gg.setup(_t(E))
    return og, gg, len(free)


def _close(a, b, tol, what, atol=0.0):
    if a is None or b is None:
        assert a is None and b is None, what
        return
    if isinstance(a, (bool, str)) or isinstance(b, (bool, str)):
        assert a == b, (what, a, b)
        return
    a, b = float(a), float(b)
    if math.isinf(a) or math.isinf(b):
        assert a == b, (what, a, b)
        return
    assert abs(a - b) <= tol * max(abs(a), abs(b)) + atol, (what, a, b)


STACKS = [dict(), dict(coarse_correction_policy="none", root_local_correction_mode="none"),
          dict(root_enrichment_mode="weighted_gap_second"),
          dict(root_local_correction_mode="node_block_line_search"),
          dict(cycle_type="w"), dict(cycle_type="w", coarse_correction_policy="none",
                                     root_local_correction_mode="none")]


@pytest.mark.parametrize("stack", STACKS)
@pytest.mark.parametrize("rho_min", [1e-3, 1e-12])
def test_cycle_trace_matches_oracle(stack, rho_min):
    og, gg, n = _pair((16, 8, 8), rho_min, **stack)
    rhs = np.random.default_rng(21).standard_normal(n)
    to = og.cycle_trace(rhs)
    tg = gg.cycle_trace(_t(rhs)).to_dict()
    vo, vg = to.pop("visits"), tg.pop("visits")
    out_o = to.pop("output")
    assert len(vo) == len(vg)
    # Rounding noise is compared only for its size, against an absolute floor
    # of 1e-12 of the cycle's |rhs| (norms) or 1e-12 (ratios, scales):
    # * the residual of an exact dense coarsest solve;
    # * a second coarse correction whose coarse rhs is zero in exact
    #   arithmetic (W cycle, no line search, exact coarsest solve: Galerkin
    #   orthogonality gives R r1 = 0), and the coarsest visit it drives;
    # * a raw line-search scale that is zero in exact arithmetic (W cycle
    #   with line search above an exact coarsest solve: the second correction
    #   repeats the first direction, to which the line-searched residual is
    #   orthogonal), whose sign and "nonpositive" note are arbitrary.
    root = to["rhs_norm"]
    small = ("ratio", "vs_", "alignment", "z_over", "scale")
    tiny = lambda x: x is not None and abs(x) <= 1e-10 * root
    for k in tg:
        _close(tg[k], to[k], 1e-8, k, 1e-12 * (1.0 if any(r in k for r in small) else root))
    for a, b in zip(vg, vo):
        assert set(a) == set(b)
        for k in ("visit_id", "level", "cycle_type", "is_coarsest", "n_coarse_corrections",
                  "coarsest_mode"):
            assert a[k] == b[k], (a["visit_id"], k)
        if tiny(b["rhs_norm"]):
            assert tiny(a["rhs_norm"]), a["visit_id"]
            continue
        second_noise = (b["second_restricted_rhs_norm"] is not None
                        and b["second_restricted_rhs_norm"] <= 1e-10 * b["restricted_rhs_norm"])
        scale_noise = any(x is not None and abs(x) <= 1e-12 for x in
                          (a["correction_scale_raw"], b["correction_scale_raw"],
                           a["second_correction_scale_raw"], b["second_correction_scale_raw"]))
        for k in a:
            if k == "notes" and scale_noise:
                strip = lambda n: [t for t in n.split("|") if t and t != "nonpositive_correction_scale"]
                assert strip(a[k]) == strip(b[k]), (a["visit_id"], a[k], b[k])
                continue
            if second_noise and k in ("second_restricted_rhs_norm", "second_coarse_correction_norm",
                                      "second_prolongated_correction_norm",
                                      "second_correction_alignment"):
                if k == "second_restricted_rhs_norm":
                    assert a[k] <= 1e-10 * a["restricted_rhs_norm"], a["visit_id"]
                continue
            sc = 1.0 if any(r in k for r in small) else root
            _close(a[k], b[k], 1e-8, (a["visit_id"], k), 1e-12 * sc)
    z = gg._last_trace_output.cpu().numpy()
    assert np.abs(z - out_o).max() <= 1e-8 * np.abs(out_o).max()
    # same kernels as the ordinary application
    assert np.array_equal(z, gg.apply(_t(rhs)).cpu().numpy())


@pytest.mark.parametrize("rho_min", [1e-3, 1e-12])
def test_kappa_eff_matches_oracle(rho_min):
    lin = dict(coarse_correction_policy="none", root_local_correction_mode="none")
    og, gg, n = _pair((16, 8, 8), rho_min, **lin)
    ko = og.estimate_kappa_eff(n_iter=25)
    kg = gg.estimate_kappa_eff(n_iter=25)
    print("kappa_eff", rho_min, kg, ko)
    assert kg == gg.kappa_eff
    assert abs(kg - ko) <= 1e-6 * ko
    og2, gg2, _ = _pair((16, 8, 8), rho_min)
    with pytest.raises(RuntimeError):
        gg2.estimate_kappa_eff()
