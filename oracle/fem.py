"""Oracle: problem set-up and fine-level operators (NumPy).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Each function cites the reference lines it restates.  Pinning (checks in
``tests/test_oracle_fem.py``):

* ``ke_hex8``            closed forms: symmetry, 6 rigid-body zero modes,
                         diag = 55/117, lambda_max = 2.5, trace.
* ``edof_hex8``          brute-force node numbering on tiny grids.
* ``generate_bc``        fixture columns (free-DOF counts of the reference
                         configurations) + explicit small cases.
* ``MatrixFreeKff``      assembled ``scipy.sparse`` K restricted to free DOFs.
* ``matvec_bf16``        the assembled K built from the bf16-rounded KE
                         applied to the bf16-rounded u (torch's own bfloat16
                         conversion), and the exact-product identity
                         KE_bf16 @ u_bf16 == float64 product of the rounded
                         operands for bf16-representable inputs.
* ``extract_diagonal``   diagonal of the assembled K_ff.
* ``NodeBlockJacobiMF``  3x3 diagonal blocks of the assembled K, inverted.
* ``element_energy``     energy identity sum_e E_e ce_e = u^T K(E) u against
                         the assembled K, and a finite-difference check of
                         d(F.u)/d rho_e = -p E'(rho) ce_e (self-adjoint).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp


# ---------------------------------------------------------------------------
# Element stiffness (pub_simp_solver.py:157-186)
# ---------------------------------------------------------------------------

_HEX_CORNERS = np.array(
    [[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
     [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)


def ke_hex8(nu: float = 0.3, E: float = 1.0) -> np.ndarray:
    """24x24 trilinear hex stiffness on [-1,1]^3, 2x2x2 Gauss rule.

    Restates pub_simp_solver.py:157-183 (isotropic D, B-matrix with engineering
    shear ordering xx,yy,zz,xy,yz,xz; DOF order 3*node+component).
    """
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[0, 0] = D[1, 1] = D[2, 2] = lam + 2 * mu
    D[3, 3] = D[4, 4] = D[5, 5] = mu
    g = 1.0 / np.sqrt(3.0)
    KE = np.zeros((24, 24))
    for r, s, t in _HEX_CORNERS * g:
        cx, cy, cz = _HEX_CORNERS[:, 0], _HEX_CORNERS[:, 1], _HEX_CORNERS[:, 2]
        # shape function derivatives N_i = (1+cx r)(1+cy s)(1+cz t)/8
        dN = 0.125 * np.stack([cx * (1 + cy * s) * (1 + cz * t),
                               cy * (1 + cx * r) * (1 + cz * t),
                               cz * (1 + cx * r) * (1 + cy * s)], axis=1)
        J = dN.T @ _HEX_CORNERS
        dxyz = dN @ np.linalg.inv(J).T
        B = np.zeros((6, 24))
        i = np.arange(8)
        B[0, 3 * i] = dxyz[:, 0]
        B[1, 3 * i + 1] = dxyz[:, 1]
        B[2, 3 * i + 2] = dxyz[:, 2]
        B[3, 3 * i] = dxyz[:, 1]
        B[3, 3 * i + 1] = dxyz[:, 0]
        B[4, 3 * i + 1] = dxyz[:, 2]
        B[4, 3 * i + 2] = dxyz[:, 1]
        B[5, 3 * i] = dxyz[:, 2]
        B[5, 3 * i + 2] = dxyz[:, 0]
        KE += B.T @ D @ B * abs(np.linalg.det(J))
    return KE


KE_UNIT_3D = ke_hex8()


def edof_hex8(nelx: int, nely: int, nelz: int) -> np.ndarray:
    """(n_elem, 24) int32 element->DOF table (pub_simp_solver.py:196-207).

    Element e = ix*nely*nelz + iy*nelz + iz; node n = ix*(nely+1)(nelz+1) +
    iy*(nelz+1) + iz; local node order (0,0,0),(1,0,0),(1,1,0),(0,1,0),
    (0,0,1),(1,0,1),(1,1,1),(0,1,1).
    """
    nny, nnz = nely + 1, nelz + 1
    ex, ey, ez = [a.ravel() for a in np.meshgrid(
        np.arange(nelx), np.arange(nely), np.arange(nelz), indexing="ij")]
    offs = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
            (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    nodes = np.stack([(ex + a) * nny * nnz + (ey + b) * nnz + (ez + c)
                      for a, b, c in offs], axis=1)
    edof = 3 * nodes[:, :, None] + np.arange(3)[None, None, :]
    return edof.reshape(len(ex), 24).astype(np.int32)


# ---------------------------------------------------------------------------
# Problem specs and boundary conditions (problem_spec.py, presets.py,
# bc_generator.py:63-251, 381-447)
# ---------------------------------------------------------------------------

@dataclass
class Support:
    kind: str            # "face" or "point"
    where: object        # face name or (x, y, z)
    constraint: str = "fixed"


@dataclass
class Load:
    kind: str            # "point", "face" (pressure) or "edge_nodes"
    where: object
    f: tuple = (0.0, 0.0, 0.0)
    magnitude: float = -1.0


@dataclass
class Spec:
    Lx: float
    Ly: float
    Lz: float
    nelx: int
    nely: int
    nelz: int
    supports: list = field(default_factory=list)
    loads: list = field(default_factory=list)
    name: str = ""


def _dof_mask(constraint: str) -> list:
    # problem_spec.py:41-52
    return {"fixed": [0, 1, 2], "pin_x": [0], "pin_y": [1], "pin_z": [2],
            "roller_x": [1, 2], "roller_y": [0, 2], "roller_z": [0, 1]}.get(
                constraint, [0, 1, 2])


def _cantilever(n, name, load="tip"):
    nelx, nely, nelz = n
    loads = ([Load("point", (2.0, 0.5, 0.25), f=(0.0, -1.0, 0.0))]
             if load == "tip" else [Load("edge_nodes", "right_bottom")])
    return Spec(2.0, 1.0, 0.5, nelx, nely, nelz,
                supports=[Support("face", "left", "fixed")], loads=loads, name=name)


def _bridge(n, name, zs):
    nelx, nely, nelz = n
    sup = [Support("point", (0.0, 0.0, z), "fixed") for z in zs]
    sup += [Support("point", (3.0, 0.0, z), "roller_x") for z in zs]
    return Spec(3.0, 1.0, 0.5, nelx, nely, nelz, supports=sup,
                loads=[Load("face", "top", magnitude=-1.0)], name=name)


def _bracket(n, name):
    # presets.py:363-369, 425-440: top (y = Ly) face fixed, point load
    # (fx, fy) = (0.3, -0.7) at (1.0, 0.0, 0.25)
    nelx, nely, nelz = n
    return Spec(1.0, 2.0, 0.5, nelx, nely, nelz, supports=[Support("face", "top", "fixed")],
                loads=[Load("point", (1.0, 0.0, 0.25), f=(0.3, -0.7, 0.0))], name=name)


PRESETS = {
    # presets.py:34-40, 137-170 (cantilever family, tip point load)
    "cantilever_3d": lambda: _cantilever((24, 12, 6), "cantilever_3d"),
    "cantilever_gpu_medium": lambda: _cantilever((80, 40, 20), "cantilever_gpu_medium"),
    "cantilever_gpu_large": lambda: _cantilever((120, 60, 30), "cantilever_gpu_large"),
    "cantilever_gpu_xlarge": lambda: _cantilever((160, 80, 40), "cantilever_gpu_xlarge"),
    "cantilever_gpu_xxlarge": lambda: _cantilever((200, 100, 50), "cantilever_gpu_xxlarge"),
    # presets.py:43-55, 265-278 (bridge family, distributed top load)
    "bridge_3d": lambda: _bridge((30, 10, 5), "bridge_3d", (0.0, 0.25, 0.5)),
    "bridge_gpu_medium": lambda: _bridge((90, 30, 15), "bridge_gpu_medium", (0.0, 0.25, 0.5)),
    # presets.py:363-369, 425-440 (bracket family)
    "bracket_3d": lambda: _bracket((12, 24, 6), "bracket_3d"),
    "bracket_gpu_medium": lambda: _bracket((30, 60, 15), "bracket_gpu_medium"),
    "bracket_gpu_500k": lambda: _bracket((80, 160, 40), "bracket_gpu_500k"),
}


def get_preset(name: str) -> Spec:
    return PRESETS[name]()


def cantilever_spec(nelx: int, nely: int, nelz: int, load: str = "tip") -> Spec:
    """Cantilever of the preset geometry at arbitrary resolution.

    ``load="tip"``: the reference's point load (2.0, 0.5, 0.25) fy=-1.
    ``load="edge"``: unit total downward load split equally over the nodes of
    the bottom edge (y=0) of the free end x=L (all z).
    """
    return _cantilever((nelx, nely, nelz), f"cantilever_{nelx}x{nely}x{nelz}_{load}", load)


def _node_coords(spec: Spec) -> np.ndarray:
    ix, iy, iz = np.meshgrid(np.arange(spec.nelx + 1), np.arange(spec.nely + 1),
                             np.arange(spec.nelz + 1), indexing="ij")
    return np.column_stack([ix.ravel() * (spec.Lx / spec.nelx),
                            iy.ravel() * (spec.Ly / spec.nely),
                            iz.ravel() * (spec.Lz / spec.nelz)])


def _nid(i, j, k, spec):
    return (i * (spec.nely + 1) + j) * (spec.nelz + 1) + k


def _face_nodes(face: str, spec: Spec) -> np.ndarray:
    # bc_generator.py:129-162 (same enumeration order)
    X, Y, Z = spec.nelx, spec.nely, spec.nelz
    if face == "left":
        return np.array([_nid(0, j, k, spec) for j in range(Y + 1) for k in range(Z + 1)])
    if face == "right":
        return np.array([_nid(X, j, k, spec) for j in range(Y + 1) for k in range(Z + 1)])
    if face == "bottom":
        return np.array([_nid(i, 0, k, spec) for i in range(X + 1) for k in range(Z + 1)])
    if face == "top":
        return np.array([_nid(i, Y, k, spec) for i in range(X + 1) for k in range(Z + 1)])
    if face == "front":
        return np.array([_nid(i, j, 0, spec) for i in range(X + 1) for j in range(Y + 1)])
    if face == "back":
        return np.array([_nid(i, j, Z, spec) for i in range(X + 1) for j in range(Y + 1)])
    raise ValueError(face)


def _nearest(coords, pt) -> int:
    return int(np.argmin(np.linalg.norm(coords - np.asarray(pt)[None, :], axis=1)))


def _tributary(ncoords, spec, face):
    # bc_generator.py:312-337 (3D branch)
    if face in ("left", "right"):
        h1, h2, c1, c2 = spec.Ly / spec.nely, spec.Lz / spec.nelz, ncoords[:, 1], ncoords[:, 2]
    elif face in ("bottom", "top"):
        h1, h2, c1, c2 = spec.Lx / spec.nelx, spec.Lz / spec.nelz, ncoords[:, 0], ncoords[:, 2]
    else:
        h1, h2, c1, c2 = spec.Lx / spec.nelx, spec.Ly / spec.nely, ncoords[:, 0], ncoords[:, 1]
    w1 = np.full(len(ncoords), h1)
    w1[np.abs(c1 - c1.min()) < h1 * 0.01] = h1 / 2
    w1[np.abs(c1 - c1.max()) < h1 * 0.01] = h1 / 2
    w2 = np.full(len(ncoords), h2)
    w2[np.abs(c2 - c2.min()) < h2 * 0.01] = h2 / 2
    w2[np.abs(c2 - c2.max()) < h2 * 0.01] = h2 / 2
    return w1 * w2


@dataclass
class BC:
    fixed_dofs: np.ndarray
    free_dofs: np.ndarray
    F: np.ndarray
    ndof: int


def generate_bc(spec: Spec) -> BC:
    """Fixed/free DOFs and load vector (bc_generator.py:381-447)."""
    coords = _node_coords(spec)
    ndof = 3 * (spec.nelx + 1) * (spec.nely + 1) * (spec.nelz + 1)
    fixed = set()
    for s in spec.supports:
        nodes = (_face_nodes(s.where, spec) if s.kind == "face"
                 else [_nearest(coords, s.where)])
        for n in nodes:
            for off in _dof_mask(s.constraint):
                fixed.add(3 * int(n) + off)
    fixed = np.sort(np.array(sorted(fixed), dtype=np.int64))
    free = np.setdiff1d(np.arange(ndof), fixed)
    F = np.zeros(ndof)
    for ld in spec.loads:
        if ld.kind == "point":
            n = _nearest(coords, ld.where)
            F[3 * n] += ld.f[0]
            F[3 * n + 1] += ld.f[1]
            F[3 * n + 2] += ld.f[2]
        elif ld.kind == "face":
            nodes = _face_nodes(ld.where, spec)
            w = _tributary(coords[nodes], spec, ld.where)
            off, sign = {"left": (0, -1.0), "right": (0, 1.0), "bottom": (1, -1.0),
                         "top": (1, 1.0), "front": (2, -1.0), "back": (2, 1.0)}[ld.where]
            for i, n in enumerate(nodes):
                F[3 * n + off] += ld.magnitude * sign * w[i]
        elif ld.kind == "edge_nodes":
            nodes = [_nid(spec.nelx, 0, k, spec) for k in range(spec.nelz + 1)]
            for n in nodes:
                F[3 * n + 1] += -1.0 / len(nodes)
        else:
            raise ValueError(ld.kind)
    F[fixed] = 0.0
    if np.max(np.abs(F[free])) < 1e-30:
        raise ValueError("Force vector is zero on all free DOFs.")
    return BC(fixed, free, F, ndof)


def assemble_K(edof: np.ndarray, E: np.ndarray, KE: np.ndarray, ndof: int) -> sp.csr_matrix:
    """Brute-force global stiffness (for checks only)."""
    rows = np.repeat(edof, 24, axis=1).ravel()
    cols = np.tile(edof, (1, 24)).ravel()
    vals = (E[:, None, None] * KE[None, :, :]).reshape(len(E), 576).ravel()
    K = sp.coo_matrix((vals, (rows, cols)), shape=(ndof, ndof)).tocsr()
    K.sum_duplicates()
    return K


# ---------------------------------------------------------------------------
# Matrix-free operator (solver_v2.py:1121-1251)
# ---------------------------------------------------------------------------

class MatrixFreeKff:
    """y_free = K_ff(E) u_free by gather -> KE @ u_e -> *E -> bincount."""

    def __init__(self, edof, KE, free, n_free, ndof):
        self._edof = np.asarray(edof)
        self._KE = np.asarray(KE, dtype=np.float64)
        self._free = np.asarray(free)
        self._n_free = int(n_free)
        self._ndof = int(ndof)
        self._n_elem, self._ke_size = self._edof.shape
        self._edof_flat = self._edof.ravel().astype(np.int64)
        self._KE_diag = self._KE.diagonal().copy()

    def matvec(self, u_free, E_e):
        # solver_v2.py:1198-1222
        dt = u_free.dtype
        u_full = np.zeros(self._ndof, dtype=dt)
        u_full[self._free] = u_free
        u_elem = u_full[self._edof]
        KE = self._KE if dt == np.float64 else self._KE.astype(dt)
        f_elem = (KE @ u_elem.T * E_e[None, :]).T
        y_full = np.bincount(self._edof_flat, weights=f_elem.ravel(),
                             minlength=self._ndof).astype(dt, copy=False)
        return y_full[self._free]

    def matvec_bf16(self, u_free, E_e):
        """The fused matvec's bf16 form (cuda_fused_matvec.py:138-246,
        343-377): u (FP32, zero on fixed DOFs) and KE (via FP32) rounded to
        bf16 to nearest even, element products KE_bf16 @ u_e accumulated in
        FP32, scaled by E_e in FP32, summed per DOF in FP32; y[free] as
        float32.  The products of bf16 operands are exact in FP32; here each
        24-term element sum and each DOF sum is formed exactly in float64
        and rounded once to float32, a result within a few FP32 ulps of any
        FP32 summation order (the reference's atomics have none fixed)."""
        from .cnn import bf16_round

        u_full = np.zeros(self._ndof, dtype=np.float32)
        u_full[self._free] = np.asarray(u_free, dtype=np.float32)
        ub = bf16_round(u_full)
        KEb = bf16_round(self._KE.astype(np.float32))
        f_elem = (ub[self._edof] @ KEb.T).astype(np.float32)
        f_elem = (f_elem * np.asarray(E_e, dtype=np.float32)[:, None]).astype(np.float32)
        y_full = np.bincount(self._edof_flat, weights=f_elem.astype(np.float64).ravel(),
                             minlength=self._ndof).astype(np.float32)
        return y_full[self._free]

    def extract_diagonal(self, E_e):
        # solver_v2.py:1243-1251
        contrib = E_e[:, None] * self._KE_diag[None, :]
        y_full = np.bincount(self._edof_flat, weights=contrib.ravel(), minlength=self._ndof)
        return y_full[self._free]


class NodeBlockJacobiMF:
    """3x3 node-block Jacobi (solver_v2.py:1268-1451)."""

    def __init__(self, edof, KE, free, n_free, ndof):
        self._n_free = int(n_free)
        self._ndof = int(ndof)
        n_nodes = ndof // 3
        self._n_nodes = n_nodes
        edof = np.asarray(edof)
        KE = np.asarray(KE)
        self._ke_blocks = np.stack([KE[3 * i:3 * i + 3, 3 * i:3 * i + 3] for i in range(8)])
        self._global_node = (edof[:, [3 * i for i in range(8)]] // 3).ravel().astype(np.int64)
        self._local_node = np.tile(np.arange(8), edof.shape[0])
        free = np.asarray(free)
        mask = np.zeros(ndof, dtype=bool)
        mask[free] = True
        node_free = mask.reshape(n_nodes, 3)
        all_free = node_free.all(axis=1)
        self._all_free_nodes = np.where(all_free)[0]
        f2f = np.full(ndof, -1, dtype=np.int64)
        f2f[free] = np.arange(len(free))
        self._all_free_local = f2f[3 * self._all_free_nodes[:, None] + np.arange(3)[None, :]]
        partial = node_free.any(axis=1) & ~all_free
        pl = []
        for n in np.where(partial)[0]:
            for d in range(3):
                if node_free[n, d]:
                    pl.append(int(f2f[3 * n + d]))
        self._partial_local = np.array(pl, dtype=np.int64)
        self._block_inv = None
        self._partial_diag_inv = None

    def update(self, E_e, scalar_diag=None):
        # solver_v2.py:1393-1423
        E_pair = np.repeat(E_e, 8)
        block_K = np.zeros((self._n_nodes, 3, 3))
        for r in range(3):
            for c in range(3):
                block_K[:, r, c] = np.bincount(
                    self._global_node, weights=E_pair * self._ke_blocks[self._local_node, r, c],
                    minlength=self._n_nodes)
        self._block_inv = np.linalg.inv(block_K[self._all_free_nodes])
        if self._partial_local.size > 0:
            if scalar_diag is not None:
                dv = scalar_diag[self._partial_local]
                self._partial_diag_inv = 1.0 / np.where(np.abs(dv) > 1e-12, dv, 1.0)
            else:
                self._partial_diag_inv = None

    def apply(self, r):
        # solver_v2.py:1429-1451
        z = np.zeros(self._n_free)
        if len(self._all_free_nodes) > 0 and self._block_inv is not None:
            rb = r[self._all_free_local].astype(np.float64)
            z[self._all_free_local] = np.einsum("nij,nj->ni", self._block_inv, rb)
        if self._partial_local.size > 0 and self._partial_diag_inv is not None:
            z[self._partial_local] = self._partial_diag_inv * r[self._partial_local].astype(np.float64)
        return z.astype(r.dtype)


def element_energy(edof: np.ndarray, KE: np.ndarray, u_full: np.ndarray) -> np.ndarray:
    """ce_e = u_e^T KE u_e (solver_v4.py:258-262: KUe = Ue @ KE, ce = sum(KUe * Ue))."""
    Ue = u_full[edof]
    return ((Ue @ KE) * Ue).sum(axis=1)


def build_components(spec: Spec, KE=None):
    """(spec, bc, free, edof, F_free, mf_op) -- the CPU half of
    run_experiments_e1_e10.py:128-179 (the GMG object is built by gmg.py)."""
    KE = KE_UNIT_3D if KE is None else KE
    bc = generate_bc(spec)
    free = bc.free_dofs.astype(np.int32)
    edof = edof_hex8(spec.nelx, spec.nely, spec.nelz)
    F_free = bc.F[free].astype(np.float64)
    mf = MatrixFreeKff(edof, KE, free, len(free), bc.ndof)
    return bc, free, edof, F_free, mf
