"""Element stiffness and element/DOF tables of the structured hex mesh.

Host-side set-up (NumPy), the same quantities as pub_simp_solver.py:157-213 of
the reference: the 24x24 trilinear hex stiffness on the [-1,1]^3 cube
(2x2x2 Gauss rule, isotropic, E=1, nu=0.3) and the element -> DOF table in the
reference numbering.  The device kernels never read the DOF table (indices are
computed from the grid); it is kept because the reference operator
constructors take it.
"""

from __future__ import annotations

import numpy as np

# local node positions in the reference ordering
HEX_NODES = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                      (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], dtype=np.int64)


def _elasticity_matrix(E: float, nu: float) -> np.ndarray:
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.diag([lam + 2 * mu] * 3 + [mu] * 3).astype(np.float64)
    D[0, 1] = D[0, 2] = D[1, 0] = D[1, 2] = D[2, 0] = D[2, 1] = lam
    return D


def _build_ke_unit_3d(nu: float = 0.3, E: float = 1.0) -> np.ndarray:
    """K_e = sum_g B_g^T D B_g |det J_g| over the 8 Gauss points +-1/sqrt(3).

    Strain order xx, yy, zz, xy, yz, xz (pub_simp_solver.py:157-183)."""
    D = _elasticity_matrix(E, nu)
    s = 2.0 * HEX_NODES - 1.0                      # corner signs, (8,3)
    gp = 1.0 / np.sqrt(3.0)
    K = np.zeros((24, 24))
    for g in s * gp:
        # dN_i/dxi_a = s_ia/8 * prod_{b != a} (1 + s_ib g_b)
        f = 1.0 + s * g[None, :]                   # (8,3)
        dN = np.empty((8, 3))
        dN[:, 0] = s[:, 0] * f[:, 1] * f[:, 2] / 8.0
        dN[:, 1] = s[:, 1] * f[:, 0] * f[:, 2] / 8.0
        dN[:, 2] = s[:, 2] * f[:, 0] * f[:, 1] / 8.0
        J = dN.T @ s
        G = dN @ np.linalg.inv(J).T                # physical gradients (8,3)
        B = np.zeros((6, 24))
        for i in range(8):
            gx, gy, gz = G[i]
            c = 3 * i
            B[0, c], B[1, c + 1], B[2, c + 2] = gx, gy, gz
            B[3, c], B[3, c + 1] = gy, gx
            B[4, c + 1], B[4, c + 2] = gz, gy
            B[5, c], B[5, c + 2] = gz, gx
        K += B.T @ D @ B * abs(np.linalg.det(J))
    return K


KE_UNIT_3D = _build_ke_unit_3d()


def _edof_table_3d(nelx: int, nely: int, nelz: int) -> np.ndarray:
    """(n_elem, 24) int32 element -> global DOF table (pub_simp_solver.py:196-207)."""
    e = np.arange(nelx * nely * nelz, dtype=np.int64)
    ex, ey, ez = e // (nely * nelz), (e // nelz) % nely, e % nelz
    sy, sx = nelz + 1, (nely + 1) * (nelz + 1)
    base = ex * sx + ey * sy + ez
    offs = HEX_NODES[:, 0] * sx + HEX_NODES[:, 1] * sy + HEX_NODES[:, 2]
    nodes = base[:, None] + offs[None, :]
    return (3 * nodes[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 24).astype(np.int32)


def _build_sparse_indices(edof: np.ndarray):
    """Row/column index pairs of the element matrices (pub_simp_solver.py:209-213)."""
    rows = np.repeat(edof, edof.shape[1], axis=1)
    cols = np.tile(edof, (1, edof.shape[1]))
    return rows.ravel().astype(np.int32), cols.ravel().astype(np.int32)


def grid_from_edof(edof: np.ndarray):
    """Recover (nelx, nely, nelz) from a structured hex DOF table and verify it.

    Raises ValueError when the table is not the structured numbering (the
    HIP operator computes indices from the grid and needs that structure)."""
    edof = np.asarray(edof)
    if edof.ndim != 2 or edof.shape[1] != 24 or edof.shape[0] < 1:
        raise ValueError("edof must be an (n_elem, 24) hex8 DOF table")
    n0 = edof[0, ::3] // 3
    sy = int(n0[3] - n0[0])          # node (0,1,0)
    sx = int(n0[1] - n0[0])          # node (1,0,0)
    nelz = sy - 1
    if nelz < 1 or sx % sy != 0:
        raise ValueError("edof is not a structured hex numbering")
    nely = sx // sy - 1
    if nely < 1 or edof.shape[0] % (nely * nelz) != 0:
        raise ValueError("edof is not a structured hex numbering")
    nelx = edof.shape[0] // (nely * nelz)
    if not np.array_equal(edof, _edof_table_3d(nelx, nely, nelz)):
        raise ValueError("edof differs from the structured hex numbering")
    return nelx, nely, nelz
