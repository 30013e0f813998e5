"""Device DOF numbering and batched element assembly through libvbh.

``assemble_model`` replaces the reference's host assembly loop
(src/solver/assembly_parallel.hpp:68-110 + SparseMatrixBuilder::build_csr,
src/core/sparse_matrix.cpp:16-150): element matrices are computed and
scattered on the GPU, and the lower CSR operator stays resident in HBM as the
backend's matrix. The returned CsrData is a handle to that device matrix; its
host arrays are fetched lazily (``fetch()``) for inspection or parity tests.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .backend import CsrData, _check
from . import model as M
from .model import Model, Mpc


class DeviceCsr(CsrData):
    """CsrData whose arrays live on the device inside a backend context."""

    def __init__(self, backend, n, nnz):
        super().__init__(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0), n, nnz, "Lower")
        self._backend = backend
        self._fetched = False

    def fetch(self):
        if not self._fetched:
            lib = L.load()
            rp = np.empty(self.n + 1, dtype=np.int32)
            ci = np.empty(self.nnz, dtype=np.int32)
            v = np.empty(self.nnz, dtype=np.float64)
            _check(lib.vbh_get_csr(self._backend._h, L.ip(rp), L.ip(ci), L.dp(v)))
            self.row_ptr, self.col_ind, self.values = rp, ci, v
            self._fetched = True
        return self


def dof_map_device(backend, model: Model):
    """GPU equation numbering; returns (eq int32[nn,6], n_free)."""
    lib = L.load()
    ids = np.ascontiguousarray(model.node_ids, dtype=np.int32)
    et = np.ascontiguousarray(model.elem_type, dtype=np.int32)
    ep = np.ascontiguousarray(model.elem_ptr, dtype=np.int32)
    cn = np.ascontiguousarray(model.conn, dtype=np.int32)
    sn = np.ascontiguousarray(model.spc_nodes, dtype=np.int32)
    sm = np.ascontiguousarray(model.spc_masks, dtype=np.int32)
    eq = np.empty((model.nn, 6), dtype=np.int32)
    nf = C.c_int(0)
    _check(lib.vbh_dof_map(backend._h, len(ids), L.ip(ids), len(et), L.ip(et), L.ip(ep), L.ip(cn),
                           len(sn), L.ip(sn), L.ip(sm), L.ip(eq), C.byref(nf)))
    return eq, nf.value


def assemble_model(backend, model: Model, eq: np.ndarray, n: int) -> DeviceCsr:
    """Assemble K on the device and install it as ``backend``'s operator."""
    lib = L.load()
    xyz = np.ascontiguousarray(model.xyz, dtype=np.float64)
    eqc = np.ascontiguousarray(eq, dtype=np.int32)
    et = np.ascontiguousarray(model.elem_type, dtype=np.int32)
    ep = np.ascontiguousarray(model.elem_ptr, dtype=np.int32)
    cn = np.ascontiguousarray(model.conn, dtype=np.int32)
    E = np.ascontiguousarray(model.E, dtype=np.float64)
    nu = np.ascontiguousarray(model.nu, dtype=np.float64)
    th = None if model.thickness is None else np.ascontiguousarray(model.thickness, dtype=np.float64)
    _check(lib.vbh_assemble(backend._h, n, model.nn, L.dp(xyz), L.ip(eqc), model.ne, L.ip(et),
                            L.ip(ep), L.ip(cn), L.dp(E), L.dp(nu),
                            L.dp(th) if th is not None else None, int(model.solid_form),
                            int(model.shell_form)))
    nn_ = C.c_int(0)
    nnz = C.c_int(0)
    _check(lib.vbh_get_csr_size(backend._h, C.byref(nn_), C.byref(nnz)))
    K = DeviceCsr(backend, nn_.value, nnz.value)
    backend._uploaded = K
    return K


def set_mpc(backend, model: Model, eq: np.ndarray, n: int, mpc) -> int:
    """MpcHandler::build (reference src/core/mpc_handler.cpp:14-164) in the
    backend context for the equations ``mpc`` (an ``Mpc`` or None to clear)
    over the numbering ``eq`` of ``n`` equations. From the next
    ``assemble_model`` on, the operator is K_red = T^T K T, load vectors are
    T^T F and solutions are expanded by T for recovery. Returns n_reduced."""
    lib = L.load()
    eqc = np.ascontiguousarray(eq, dtype=np.int32)
    nr = C.c_int(0)
    if mpc is None or mpc.n == 0:
        _check(lib.vbh_set_mpc(backend._h, model.nn, L.ip(eqc), int(n), 0, None, None, None, None,
                               C.byref(nr)))
        return nr.value
    ptr = np.ascontiguousarray(mpc.ptr, dtype=np.int32)
    node = np.ascontiguousarray(mpc.node, dtype=np.int32)
    dof = np.ascontiguousarray(mpc.dof, dtype=np.int32)
    coeff = np.ascontiguousarray(mpc.coeff, dtype=np.float64)
    _check(lib.vbh_set_mpc(backend._h, model.nn, L.ip(eqc), int(n), mpc.n, L.ip(ptr), L.ip(node),
                           L.ip(dof), L.dp(coeff), C.byref(nr)))
    return nr.value


def mpc_stats(backend) -> L.vbh_mpc_stats:
    s = L.vbh_mpc_stats()
    _check(L.load().vbh_mpc_info(backend._h, C.byref(s)))
    return s


def mpc_transform(backend):
    """T of the elimination set by ``set_mpc`` as CSR arrays over the full
    equations: (t_ptr, t_col, t_val)."""
    st = mpc_stats(backend)
    ptr = np.empty(st.n_full + 1, np.int32)
    col = np.empty(max(st.t_nnz, 1), np.int32)
    val = np.empty(max(st.t_nnz, 1))
    _check(L.load().vbh_mpc_transform(backend._h, L.ip(ptr), L.ip(col), L.dp(val)))
    return ptr, col[:st.t_nnz], val[:st.t_nnz]


def mpc_expand(backend, u_red, n_full: int) -> np.ndarray:
    """u_full = T u_red over the mesh of the last ``assemble_model``
    (MpcHandler::recover_dependent_dofs, mpc_handler.cpp:312-330)."""
    u = np.ascontiguousarray(u_red, dtype=np.float64)
    out = np.empty(n_full)
    _check(L.load().vbh_mpc_expand(backend._h, L.dp(u), L.dp(out)))
    return out


def analysis_numbering(backend, model: Model):
    """Equation numbering and MPC elimination of the linear-static analysis
    (build_analysis_dof_map + build_analysis_mpc_system + MpcHandler::build,
    reference src/solver/analysis_support.cpp:176-282): returns (eq, n_full,
    n_reduced) with the elimination installed for the next ``assemble_model``.
    Without MPCs or CD frames n_reduced == n_full."""
    eq, n = M.analysis_dof_map(model)
    eq, n, mpc = M.analysis_mpc_system(model, eq, n)
    return eq, n, set_mpc(backend, model, eq, n, mpc)


def assembly_stats(backend) -> L.vbh_assembly_stats:
    s = L.vbh_assembly_stats()
    _check(L.load().vbh_last_assembly_stats(backend._h, C.byref(s)))
    return s


def element_sensitivity(backend, model: Model, param: int, u, v) -> np.ndarray:
    """out[e] = -u_e^T (dK_e/dp_e) v_e over the mesh of the last
    ``assemble_model`` on ``backend``: p = E (param 0) or CQUAD4 thickness
    (param 1). Runs the GPU adjoint kernel (no reference counterpart)."""
    lib = L.load()
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty(model.ne, dtype=np.float64)
    _check(lib.vbh_element_sensitivity(backend._h, int(param), L.dp(u), L.dp(v), L.dp(out)))
    return out


def compliance_gradients(backend, model: Model, eq, n, F, tip_dof=None):
    """Compliance C = F^T u and dC/dE_e; with ``tip_dof`` also u[tip_dof] and
    d u[tip_dof] / dE_e from the adjoint solve K lambda = e_tip. Both load
    vectors go through one block solve. Returns (C, dC, u_tip, du_tip)."""
    K = assemble_model(backend, model, eq, n)
    if tip_dof is None:
        u = backend.solve(K, F)
        return float(F @ u), element_sensitivity(backend, model, 0, u, u), None, None
    R = np.zeros((n, 2))
    R[:, 0] = F
    R[tip_dof, 1] = 1.0
    U = backend.solve_block(K, R)
    u, lam = U[:, 0], U[:, 1]
    return (float(F @ u), element_sensitivity(backend, model, 0, u, u), float(u[tip_dof]),
            element_sensitivity(backend, model, 0, u, lam))


def load_arrays(model: Model, case):
    """The C-ABI load description of one load case: nodal [nn, 6] (or None),
    PLOAD4 faces (corner node indices in outward order, [nface, 4]) with
    their corner pressures, GRAV vector (or None), element densities."""
    from .model import hexa_pload4_faces
    nodal = np.ascontiguousarray(case.nodal, dtype=np.float64) if case.nodal is not None else None
    fn = fp = None
    if case.pload4 is not None and len(case.pload4.elem):
        pl = case.pload4
        faces = hexa_pload4_faces(model, pl)
        fn = np.ascontiguousarray(np.take_along_axis(model.conn2d()[pl.elem], faces, axis=1),
                                  dtype=np.int32)
        fp = np.ascontiguousarray(pl.p, dtype=np.float64)
    g = None if case.grav is None else np.ascontiguousarray(case.grav, dtype=np.float64)
    rho = np.ascontiguousarray(model.rho, dtype=np.float64)
    return nodal, fn, fp, g, rho


def load_args(arrays):
    """ctypes arguments (nodal, nface, face_nodes, face_p, grav, rho) of
    ``load_arrays`` output; the arrays must outlive the call."""
    nodal, fn, fp, g, rho = arrays
    nface = 0 if fn is None else len(fn)
    return (L.dp(nodal) if nodal is not None else None, nface,
            L.ip(fn) if nface else None, L.dp(fp) if nface else None,
            L.dp(g) if g is not None else None, L.dp(rho))


def assemble_rhs_device(backend, model: Model, eq, n, case, F_device=None):
    """Load vector of one load case on the GPU (FORCE/MOMENT, PLOAD4 on CHEXA8
    faces, GRAV with consistent masses) over the mesh of the last
    ``assemble_model``. Returns F (length n: the operator's equations,
    n_reduced under an MPC elimination) on the host, or writes the device
    buffer ``F_device`` (an HBM address) when given."""
    lib = L.load()
    arrays = load_arrays(model, case)
    args = (backend._h, *load_args(arrays))
    if F_device is not None:
        _check(lib.vbh_assemble_rhs_device(*args, C.c_void_p(F_device)))
        return None
    F = np.empty(n, dtype=np.float64)
    _check(lib.vbh_assemble_rhs(*args, L.dp(F)))
    return F
