"""Pins for the oracle's fine-level pieces: closed forms of the hex8 stiffness,
brute-force DOF numbering, the assembled stiffness, and the reference's own
CPU eigenvalue table (direct_floor_conditioning.csv: eigsh on the 24x12x6
cantilever free-DOF matrix)."""

import csv
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import eigsh
import scipy.sparse.linalg as spla

from oracle import fem


def test_ke_closed_forms():
    KE = fem.KE_UNIT_3D
    assert KE.shape == (24, 24)
    assert np.allclose(KE, KE.T, atol=1e-15)
    ev = np.linalg.eigvalsh(KE)
    assert np.sum(np.abs(ev) < 1e-12) == 6            # rigid-body null space
    assert abs(ev[-1] - 2.5) < 1e-12
    assert abs(ev[6] - 5.0 / 39.0) < 1e-12
    assert np.allclose(np.diag(KE), 55.0 / 117.0, atol=1e-15)
    # rigid translations and rotations are in the null space
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=float)
    for t in np.eye(3):
        assert np.abs(KE @ np.tile(t, 8)).max() < 1e-13
    rot = np.zeros(24)
    rot[0::3], rot[1::3] = -X[:, 1], X[:, 0]           # rotation about z
    assert np.abs(KE @ rot).max() < 1e-13


def test_ke_scales_with_modulus():
    assert np.allclose(fem.ke_hex8(0.3, 2.0), 2.0 * fem.KE_UNIT_3D, atol=1e-14)


def test_edof_brute_force():
    nx, ny, nz = 3, 2, 2
    ed = fem.edof_hex8(nx, ny, nz)
    pos = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    e = 0
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                for k, (a, b, c) in enumerate(pos):
                    n = ((ix + a) * (ny + 1) + iy + b) * (nz + 1) + iz + c
                    assert list(ed[e, 3 * k:3 * k + 3]) == [3 * n, 3 * n + 1, 3 * n + 2]
                e += 1


@pytest.mark.parametrize("dims", [(4, 2, 2), (6, 4, 3), (5, 3, 4)])
@pytest.mark.parametrize("preset", ["tip", "bridge"])
def test_matvec_diag_against_assembled(dims, preset):
    if preset == "tip":
        spec = fem.cantilever_spec(*dims)
    else:
        spec = fem._bridge(dims, "b", (0.0, 0.25, 0.5))
    bc, free, edof, F, mf = fem.build_components(spec)
    rng = np.random.default_rng(1)
    E = rng.random(np.prod(dims)) + 1e-3
    K = fem.assemble_K(edof, E, fem.KE_UNIT_3D, bc.ndof)[free][:, free]
    u = rng.standard_normal(len(free))
    y = mf.matvec(u, E)
    assert np.abs(y - K @ u).max() <= 1e-13 * np.abs(K @ u).max()
    assert np.abs(mf.extract_diagonal(E) - K.diagonal()).max() <= 1e-14
    # FP32 path keeps FP32 element products (reference: KE cast, scatter in FP64)
    y32 = mf.matvec(u.astype(np.float32), E.astype(np.float32))
    assert y32.dtype == np.float32
    assert np.abs(y32 - K @ u).max() <= 2e-5 * np.abs(K @ u).max()


@pytest.mark.parametrize("dims", [(6, 3, 3), (5, 4, 3)])
def test_matvec_bf16_against_assembled_bf16_K(dims):
    """The bf16 fused form equals the assembled K built from bf16-rounded KE
    applied to the bf16-rounded u, with the rounding done by torch's own
    float32 -> bfloat16 conversion (an independent implementation), to FP32
    rounding; fixed DOFs enter as zero.  It differs from the FP64 operator by
    the operand rounding (about 2^-9 relative), which a dropped rounding step
    would not."""
    import torch

    spec = fem.cantilever_spec(*dims)
    bc, free, edof, F, mf = fem.build_components(spec)
    rng = np.random.default_rng(4)
    E = np.where(rng.random(np.prod(dims)) < 0.5, 1.0, 1e-3).astype(np.float32)
    u = rng.standard_normal(len(free)).astype(np.float32)

    def bf(a):
        return torch.as_tensor(np.asarray(a, dtype=np.float32)).to(torch.bfloat16).to(
            torch.float64).numpy()

    u_full = np.zeros(bc.ndof)
    u_full[free] = bf(u)
    Kb = fem.assemble_K(edof, E.astype(np.float64), bf(fem.KE_UNIT_3D), bc.ndof)
    ref = (Kb @ u_full)[free]
    y = mf.matvec_bf16(u, E)
    assert y.dtype == np.float32
    assert np.abs(y - ref).max() <= 4e-7 * np.abs(ref).max()
    y64 = mf.matvec(u.astype(np.float64), E.astype(np.float64))
    gap = np.abs(y - y64).max() / np.abs(y64).max()
    assert 1e-4 <= gap <= 2e-2


def test_nodeblock_against_assembled_blocks():
    spec = fem._bridge((6, 3, 3), "b", (0.0, 0.25, 0.5))
    bc, free, edof, F, mf = fem.build_components(spec)
    E = np.random.default_rng(2).random(54) + 0.01
    Kfull = fem.assemble_K(edof, E, fem.KE_UNIT_3D, bc.ndof).toarray()
    nb = fem.NodeBlockJacobiMF(edof, fem.KE_UNIT_3D, free, len(free), bc.ndof)
    d = mf.extract_diagonal(E)
    nb.update(E, scalar_diag=d)
    r = np.random.default_rng(3).standard_normal(len(free))
    z = nb.apply(r)
    f2f = {int(g): i for i, g in enumerate(free)}
    mask = np.zeros(bc.ndof, bool)
    mask[free] = True
    for n in range(bc.ndof // 3):
        dofs = [3 * n + c for c in range(3)]
        fr = [f2f[g] for g in dofs if mask[g]]
        if len(fr) == 3:
            B = Kfull[np.ix_(dofs, dofs)]
            assert np.allclose(z[fr], np.linalg.solve(B, r[fr]), rtol=1e-12, atol=0)
        elif fr:
            for g in dofs:
                if mask[g]:
                    i = f2f[g]
                    assert np.isclose(z[i], r[i] / Kfull[g, g], rtol=1e-13)


def test_bc_counts_of_reference_configurations():
    # free-DOF counts of the reference presets (ndof minus the clamped face)
    for name, nfree in [("cantilever_gpu_medium", 206640), ("cantilever_3d", 6552)]:
        bc = fem.generate_bc(fem.get_preset(name))
        assert len(bc.free_dofs) == nfree
        assert np.count_nonzero(bc.F) == 1 and bc.F.sum() == -1.0
    spec = fem.get_preset("bridge_gpu_medium")
    bc = fem.generate_bc(spec)
    assert len(bc.fixed_dofs) == 3 * 3 + 3 * 2      # 3 pinned nodes + 3 rollers
    assert abs(bc.F.sum() - (-1.0 * 3.0 * 0.5)) < 1e-12  # pressure x top-face area
    cfg = fem.generate_bc(fem.cantilever_spec(40, 20, 20, load="edge"))
    assert len(cfg.free_dofs) == 52920 and abs(cfg.F.sum() + 1.0) < 1e-15


def _conditioning_rows(golden_dir):
    with open(os.path.join(golden_dir, "direct_floor_conditioning.csv")) as f:
        return list(csv.DictReader(f))


@pytest.mark.parametrize("case", [(7, "0.1", "1e-12"), (13, "0.35", "0.001"), (19, "0.2", "0.01")])
def test_assembled_operator_matches_reference_eigs(golden_dir, case):
    """lambda_max of K_ff (eigsh, reference CPU table) pins KE, numbering,
    BCs and the SIMP modulus together; lambda_min pins the floor physics."""
    seed, p, rho = case
    rows = [r for r in _conditioning_rows(golden_dir)
            if int(r["seed"]) == seed and r["solid_probability"] == p and float(r["rho_min"]) == float(rho)]
    assert len(rows) == 1
    row = rows[0]
    spec = fem.get_preset("cantilever_3d")
    bc, free, edof, F, mf = fem.build_components(spec)
    solid = np.random.default_rng(seed).random(24 * 12 * 6) < float(p)
    rho_min = float(rho)
    E = rho_min + (1.0 - rho_min) * np.where(solid, 1.0, rho_min) ** 4.5
    K = fem.assemble_K(edof, E, fem.KE_UNIT_3D, bc.ndof)[free][:, free].tocsc()
    lmax = float(eigsh(K, k=1, which="LA", return_eigenvectors=False)[0])
    assert abs(lmax - float(row["lambda_max"])) <= 1e-9 * lmax
    lmin = float(eigsh(K, k=1, sigma=0.0, which="LM", return_eigenvectors=False)[0])
    assert abs(lmin - float(row["lambda_min"])) <= 1e-6 * abs(float(row["lambda_min"])) + 1e-18
    assert int(row["n_free"]) == len(free)


def test_element_energy_identity_and_sensitivity():
    """sum_e E_e u_e^T KE u_e = u^T K(E) u for any u (assembled K), and for
    u = K^-1 F the SIMP sensitivity -p (E0-Emin) rho^(p-1) ce_e equals the
    finite-difference derivative of the compliance F.u (self-adjoint)."""
    spec = fem.cantilever_spec(6, 3, 3)
    bc = fem.generate_bc(spec)
    edof = fem.edof_hex8(6, 3, 3)
    rng = np.random.default_rng(8)
    rho = rng.uniform(0.2, 1.0, edof.shape[0])
    p, Emin = 3.0, 1e-3
    E = Emin + (1 - Emin) * rho ** p
    u = rng.standard_normal(bc.ndof)
    K = fem.assemble_K(edof, E, fem.KE_UNIT_3D, bc.ndof)
    ce = fem.element_energy(edof, fem.KE_UNIT_3D, u)
    assert abs((E * ce).sum() - u @ (K @ u)) <= 1e-12 * abs(u @ (K @ u))
    free = bc.free_dofs

    def compliance(r):
        Ee = Emin + (1 - Emin) * r ** p
        Kf = fem.assemble_K(edof, Ee, fem.KE_UNIT_3D, bc.ndof)[free][:, free].tocsc()
        uf = spla.spsolve(Kf, bc.F[free])
        return float(bc.F[free] @ uf), uf

    c0, uf = compliance(rho)
    full = np.zeros(bc.ndof)
    full[free] = uf
    dc = -p * (1 - Emin) * rho ** (p - 1) * fem.element_energy(edof, fem.KE_UNIT_3D, full)
    for e in (0, 17, 40):
        h = 1e-6
        r1 = rho.copy()
        r1[e] += h
        r2 = rho.copy()
        r2[e] -= h
        fd = (compliance(r1)[0] - compliance(r2)[0]) / (2 * h)
        assert abs(fd - dc[e]) <= 1e-5 * abs(dc).max(), (e, fd, dc[e])
