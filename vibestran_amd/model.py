"""Mesh/model containers, DOF numbering and RHS assembly for the hot path.

Data layout (host side, handed to libvbh as flat arrays):
  node_ids  int32[nn]       external grid IDs (any order)
  xyz       float64[nn,3]   basic coordinates
  elem_type int32[ne]       VBH_CHEXA8 / VBH_CPENTA6 / VBH_CTETRA4 / VBH_CQUAD4
  elem_ptr  int32[ne+1]     offsets into conn
  conn      int32[...]      0-based node indices, reference node order
  E, nu, rho float64[ne]    per-element MAT1 data
  thickness float64[ne]     PSHELL T (shells only)
  elem_ids  int32[ne]       external element IDs (optional)
  mpc       Mpc             MPC equations of the analysis's MPC set (optional)
  rigid_mpc Mpc             RBE2 / RBE3 elements as MPC equations (optional)
  rigid_active int32[nn]    DOF mask the rigid elements activate (optional)
  cd        int32[nn]       GRID CD frame per node, 0 = basic (optional)
  cd_rot    float64[nn,3,3] rotation of the node's CD frame at the node:
                            columns are the frame's unit vectors in basic
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

CHEXA8, CPENTA6, CTETRA4, CQUAD4 = 1, 2, 3, 4
NODES_PER = {CHEXA8: 8, CPENTA6: 6, CTETRA4: 4, CQUAD4: 4}
SHELL_TYPES = (CQUAD4,)


@dataclass
class Pload4:
    """PLOAD4 on solid element faces: element index, G1/G34 as element-local
    corner indices (-1 = default face), and four corner pressures."""

    elem: np.ndarray
    g1: np.ndarray
    g34: np.ndarray
    p: np.ndarray  # [m,4]


@dataclass
class LoadCase:
    """One SUBCASE's loads: nodal FORCE/MOMENT (6 components per node) plus
    PLOAD4 faces and a GRAV acceleration vector."""

    nodal: np.ndarray                # [nn,6]
    pload4: Optional[Pload4] = None
    grav: Optional[np.ndarray] = None  # [3] acceleration in basic


@dataclass
class Mpc:
    """Multi-point constraint equations sum_t coeff_t u(node_t, dof_t) = 0:
    equation k holds the terms ptr[k] .. ptr[k+1]-1 (node as a 0-based node
    index, dof 1..6)."""

    ptr: np.ndarray
    node: np.ndarray
    dof: np.ndarray
    coeff: np.ndarray

    @property
    def n(self) -> int:
        return len(self.ptr) - 1

    @staticmethod
    def from_equations(eqs) -> "Mpc":
        """From a list of equations, each a list of (node index, dof, coeff)."""
        ptr = np.zeros(len(eqs) + 1, np.int32)
        node, dof, coeff = [], [], []
        for k, e in enumerate(eqs):
            for g, d, c in e:
                node.append(g)
                dof.append(d)
                coeff.append(c)
            ptr[k + 1] = len(node)
        return Mpc(ptr=ptr, node=np.asarray(node, np.int32), dof=np.asarray(dof, np.int32),
                   coeff=np.asarray(coeff, np.float64))

    def equations(self):
        return [[(int(self.node[t]), int(self.dof[t]), float(self.coeff[t]))
                 for t in range(self.ptr[k], self.ptr[k + 1])] for k in range(self.n)]


@dataclass
class Model:
    node_ids: np.ndarray
    xyz: np.ndarray
    elem_type: np.ndarray
    elem_ptr: np.ndarray
    conn: np.ndarray
    E: np.ndarray
    nu: np.ndarray
    rho: np.ndarray
    thickness: Optional[np.ndarray] = None
    solid_form: int = 0   # 0 EAS (default), 1 SRI
    shell_form: int = 0   # 0 MITC4+ (default), 1 Mindlin
    spc_nodes: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    spc_masks: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    loads: List[LoadCase] = field(default_factory=list)
    meta: dict = field(default_factory=dict)
    elem_ids: Optional[np.ndarray] = None  # external element IDs (default 1..ne)
    subcase_ids: Optional[List[int]] = None  # SUBCASE IDs of `loads` (default 1..)
    mpc: Optional[Mpc] = None
    rigid_mpc: Optional[Mpc] = None
    rigid_active: Optional[np.ndarray] = None
    cd: Optional[np.ndarray] = None
    cd_rot: Optional[np.ndarray] = None

    def element_ids(self) -> np.ndarray:
        if self.elem_ids is not None:
            return np.asarray(self.elem_ids, dtype=np.int32)
        return np.arange(1, self.ne + 1, dtype=np.int32)

    def subcase_id(self, k: int) -> int:
        return int(self.subcase_ids[k]) if self.subcase_ids is not None else k + 1

    @property
    def nn(self) -> int:
        return len(self.node_ids)

    @property
    def ne(self) -> int:
        return len(self.elem_type)

    def homogeneous(self) -> Optional[int]:
        t = np.unique(self.elem_type)
        return int(t[0]) if len(t) == 1 else None

    def conn2d(self) -> np.ndarray:
        t = self.homogeneous()
        if t is None:
            raise ValueError("mixed element types")
        return self.conn.reshape(self.ne, NODES_PER[t])


def analysis_dof_map(model: Model):
    """Equation numbering of build_analysis_dof_map
    (reference src/solver/analysis_support.cpp:176-214, dof_map.cpp:10-97):
    6 slots per node in sorted node-ID order; inactive slots (solid-only nodes
    lose R1-R3, mark_element_dofs 39-68) and SPC'd slots removed with the
    remaining indices shifted down. Equivalent closed form used here: eq is
    the rank of each free (node ID, dof) pair in lexicographic order. Rigid
    elements activate their components (mark_rigid_element_dofs, 70-95), MPC
    terms their DOFs (mark_explicit_mpc_dofs, 97-105); SPCs on nodes with a
    CD frame are left to analysis_mpc_system (206-207).

    Returns (eq int32[nn,6] with -1 for constrained, n_free)."""
    nn = model.nn
    active = np.zeros((nn, 6), dtype=bool)
    if model.rigid_active is not None:
        active |= ((np.asarray(model.rigid_active)[:, None] >> np.arange(6)[None, :]) & 1).astype(bool)
    if model.mpc is not None and model.mpc.n:
        active[model.mpc.node, model.mpc.dof - 1] = True
    for t in np.unique(model.elem_type):
        sel = np.nonzero(model.elem_type == t)[0]
        k = NODES_PER[int(t)]
        idx = (model.elem_ptr[sel][:, None] + np.arange(k)[None, :]).reshape(-1)
        nodes = model.conn[idx]
        if int(t) in SHELL_TYPES:
            active[nodes, :] = True
        else:
            active[nodes, 0:3] = True
    spc = _spc_flags(model)
    if model.cd is not None:
        spc[np.asarray(model.cd) != 0] = False
    return _number(model, active & ~spc)


def _spc_flags(model: Model) -> np.ndarray:
    spc = np.zeros((model.nn, 6), dtype=bool)
    if len(model.spc_nodes):
        bits = (model.spc_masks[:, None] >> np.arange(6)[None, :]) & 1
        np.logical_or.at(spc, model.spc_nodes, bits.astype(bool))
    return spc


def _number(model: Model, free: np.ndarray):
    order = np.argsort(model.node_ids, kind="stable")
    flat = free[order].reshape(-1)
    rank = np.cumsum(flat, dtype=np.int64) - 1
    eq = np.full((model.nn, 6), -1, dtype=np.int32)
    eq[order] = np.where(flat, rank, -1).reshape(model.nn, 6).astype(np.int32)
    return eq, int(flat.sum())


def _reduce_3x3(node: int, T3: np.ndarray, cd_dofs, base: int, mpcs, direct):
    """reduce_3x3_constraints (reference src/solver/analysis_support.cpp:107-171):
    the constrained local axes of a CD frame (rows of A, basic components)
    reduced by Gaussian elimination with column pivoting; a row left with one
    nonzero is a direct constraint of that basic DOF, others become MPC
    equations on the node's basic DOFs."""
    count = len(cd_dofs)
    if count == 0:
        return
    if count == 3:
        direct.extend((node, base + j) for j in range(3))
        return
    A = np.zeros((3, 3))
    perm = [0, 1, 2]
    for i in range(count):
        for j in range(3):
            A[i, j] = T3[j, cd_dofs[i]]
    for row in range(count):
        best_col, best = row, abs(A[row, row])
        for col in range(row + 1, 3):
            if abs(A[row, col]) > best:
                best, best_col = abs(A[row, col]), col
        if best_col != row:
            A[:count, [row, best_col]] = A[:count, [best_col, row]]
            perm[row], perm[best_col] = perm[best_col], perm[row]
        if abs(A[row, row]) < 1e-14:
            continue
        for i in range(row + 1, count):
            factor = A[i, row] / A[row, row]
            for j in range(row, 3):
                A[i, j] -= factor * A[row, j]
    for row in range(count):
        if abs(A[row, row]) < 1e-14:
            continue
        cols = [c for c in range(row, 3) if abs(A[row, c]) > 1e-14]
        if len(cols) == 1:
            direct.append((node, base + perm[row]))
            continue
        mpcs.append([(node, base + perm[c] + 1, float(A[row, c])) for c in cols])


def analysis_mpc_system(model: Model, eq: np.ndarray, n: int):
    """build_analysis_mpc_system (reference src/solver/analysis_support.cpp:217-282)
    up to MpcHandler::build: SPCs on nodes with a CD frame turn into direct
    constraints of basic DOFs (removed from the numbering, 263) or MPC
    equations, ahead of the rigid elements' equations (266-269) and the
    model's explicit MPC equations (270-275).

    Returns (eq, n, Mpc or None): the numbering MpcHandler::build starts from
    and every equation it eliminates."""
    eqs = []
    direct = []
    if model.cd is not None and np.any(np.asarray(model.cd) != 0):
        spc = _spc_flags(model)
        order = np.argsort(model.node_ids, kind="stable")
        for i in order:
            if model.cd[i] == 0 or not spc[i].any():
                continue
            T3 = np.asarray(model.cd_rot[i])
            trans = [d for d in range(3) if spc[i, d] and eq[i, d] >= 0]
            _reduce_3x3(int(i), T3, trans, 0, eqs, direct)
            rot = [d for d in range(3) if spc[i, d + 3] and eq[i, d + 3] >= 0]
            _reduce_3x3(int(i), T3, rot, 3, eqs, direct)
    if direct:
        free = eq >= 0
        for g, d in direct:
            free[g, d] = False
        eq, n = _number(model, free)
    if model.rigid_mpc is not None:
        eqs.extend(model.rigid_mpc.equations())
    if model.mpc is not None:
        eqs.extend(model.mpc.equations())
    return eq, n, (Mpc.from_equations(eqs) if eqs else None)


def to_output_frames(model: Model, disp: np.ndarray) -> np.ndarray:
    """Node displacements [nn, 6] from the basic frame into each node's CD
    frame, translations and rotations alike (v_cd = T3^T v_basic, reference
    src/solver/linear_static.cpp:591-612)."""
    out = np.array(disp, dtype=np.float64, copy=True)
    if model.cd is None:
        return out
    sel = np.nonzero(np.asarray(model.cd) != 0)[0]
    if len(sel):
        Tt = np.transpose(np.asarray(model.cd_rot)[sel], (0, 2, 1))
        out[sel, 0:3] = np.einsum("nij,nj->ni", Tt, out[sel, 0:3])
        out[sel, 3:6] = np.einsum("nij,nj->ni", Tt, out[sel, 3:6])
    return out


def element_gdofs(model: Model, eq: np.ndarray) -> np.ndarray:
    """global_dof_indices for a homogeneous mesh: solids [n0.T1..T3, n1...]
    (solid_elements.cpp:29-42), shells all six per node (cquad4.cpp:855-864)."""
    t = model.homogeneous()
    c = model.conn2d()
    if t in SHELL_TYPES:
        return eq[c].reshape(model.ne, -1)
    return eq[c][:, :, 0:3].reshape(model.ne, -1)


def quad_face_forces(face_xyz: np.ndarray, pressures: np.ndarray, normal_sign: float) -> np.ndarray:
    """2x2 Gauss consistent nodal forces of a 4-node face load
    (integrate_quad_surface_load, reference linear_static_pressure.cpp:154-196).
    face_xyz [m,4,3], pressures [m,4] -> forces [m,4,3]."""
    g = 1.0 / np.sqrt(3.0)
    out = np.zeros(face_xyz.shape, dtype=np.float64)
    for xi in (-g, g):
        for eta in (-g, g):
            N = 0.25 * np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta),
                                 (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)])
            dxi = 0.25 * np.array([-(1 - eta), (1 - eta), (1 + eta), -(1 + eta)])
            deta = 0.25 * np.array([-(1 - xi), -(1 + xi), (1 + xi), (1 - xi)])
            a = np.einsum("k,mkd->md", dxi, face_xyz)
            b = np.einsum("k,mkd->md", deta, face_xyz)
            p = pressures @ N
            av = np.cross(a, b)
            t = (normal_sign * p)[:, None] * av
            out += t[:, None, :] * N[None, :, None]
    return out


_HEX_FACES = np.array([[0, 1, 2, 3], [4, 5, 6, 7], [0, 1, 5, 4],
                       [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7]])


def hexa_pload4_faces(model: Model, pl: Pload4) -> np.ndarray:
    """Select and orient CHEXA8 faces like select_solid_face / orient_face_outward
    (reference linear_static_pressure.cpp:420-517). Returns local corner
    indices [m,4]."""
    c = model.conn2d()[pl.elem]
    X = model.xyz[c]                                   # [m,8,3]
    m = len(pl.elem)
    out = np.zeros((m, 4), dtype=np.int64)
    cen = X.mean(axis=1)
    for i in range(m):
        cand = range(6)
        chosen = None
        for fi in cand:
            f = _HEX_FACES[fi]
            if pl.g1[i] >= 0 or pl.g34[i] >= 0:
                if pl.g1[i] not in f or pl.g34[i] not in f:
                    continue
            nrm = np.cross(X[i, f[1]] - X[i, f[0]], X[i, f[2]] - X[i, f[0]])
            hint = cen[i] - X[i, f].mean(axis=0)
            o = f[[0, 3, 2, 1]] if nrm @ hint > 0.0 else f.copy()
            if pl.g1[i] < 0 and pl.g34[i] < 0:
                chosen = o
                break
            pos = int(np.nonzero(o == pl.g1[i])[0][0])
            rot = np.roll(o, -pos)
            if rot[2] == pl.g34[i]:
                chosen = rot
                break
        if chosen is None:
            raise ValueError(f"PLOAD4 on element index {pl.elem[i]} could not resolve face")
        out[i] = chosen
    return out


def assemble_rhs(model: Model, eq: np.ndarray, n: int, case: LoadCase) -> np.ndarray:
    """Global force vector of one load case: FORCE/MOMENT components
    (apply_point_loads, linear_static.cpp:344-398) plus PLOAD4 on CHEXA8 faces
    (apply_pressure_loads, linear_static_pressure.cpp:606-793)."""
    F = np.zeros(n, dtype=np.float64)
    nodal = np.array(case.nodal, dtype=np.float64, copy=True)
    if case.pload4 is not None and len(case.pload4.elem):
        pl = case.pload4
        faces = hexa_pload4_faces(model, pl)
        c = model.conn2d()[pl.elem]
        nodes = np.take_along_axis(c, faces, axis=1)       # [m,4]
        f = quad_face_forces(model.xyz[nodes], np.asarray(pl.p, dtype=np.float64), -1.0)
        np.add.at(nodal[:, 0:3], nodes.reshape(-1), f.reshape(-1, 3))
    if case.grav is not None:
        raise NotImplementedError("GRAV needs the consistent-mass RHS kernel")
    mask = eq >= 0
    np.add.at(F, eq[mask], nodal[mask])
    return F
