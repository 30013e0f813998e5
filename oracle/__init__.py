"""CPU FP64 oracle for the vibestran linear-static hot path.

TEST INFRASTRUCTURE ONLY. This package restates, on the CPU, the reference's
DOF numbering, element stiffness, lower-CSR assembly, symmetric matvec and
Jacobi-PCG (see ``oracle/src/oracle.cpp`` for the file:line of each
restated function). Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker.
The product path (``vibestran_amd``) never imports it.

Pinning status of each function is recorded in ``oracle/src/oracle.cpp``'s
header and checked by ``tests/test_oracle_*.py``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")

CHEXA8, CPENTA6, CTETRA4, CQUAD4 = 1, 2, 3, 4


class OracleError(RuntimeError):
    """Raised where the reference throws SolverError."""


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_SO):
        build()
    lib = C.CDLL(_SO)
    P = C.c_void_p
    I, D = C.c_int, C.c_double
    IP, DP = C.POINTER(C.c_int), C.POINTER(C.c_double)
    sig = {
        "orc_last_error": (C.c_char_p, []),
        "orc_dofmap_build": (P, [I, IP, I]),
        "orc_dofmap_free": (None, [P]),
        "orc_dofmap_constrain": (I, [P, I, I]),
        "orc_dofmap_constrain_batch": (I, [P, I, IP, IP]),
        "orc_dofmap_eq_index": (I, [P, I, I]),
        "orc_dofmap_num_free": (I, [P]),
        "orc_dofmap_num_total": (I, [P]),
        "orc_analysis_dof_map": (I, [I, IP, I, IP, IP, IP, I, IP, IP, IP]),
        "orc_hexa8_ke_sri": (I, [DP, D, D, DP]),
        "orc_hexa8_ke_eas": (I, [DP, D, D, DP]),
        "orc_tetra4_ke": (I, [DP, D, D, DP]),
        "orc_penta6_ke_eas": (I, [DP, D, D, DP]),
        "orc_hexa8_me": (I, [DP, D, DP]),
        "orc_tetra4_me": (I, [DP, D, DP]),
        "orc_penta6_me": (I, [DP, D, DP]),
        "orc_builder_create": (P, [I]),
        "orc_builder_free": (None, [P]),
        "orc_builder_add": (None, [P, I, I, D]),
        "orc_builder_add_element": (I, [P, I, IP, DP]),
        "orc_builder_num_triplets": (C.c_longlong, [P]),
        "orc_builder_build_csr": (P, [P]),
        "orc_csr_free": (None, [P]),
        "orc_csr_n": (I, [P]),
        "orc_csr_nnz": (I, [P]),
        "orc_csr_is_lower": (I, [P]),
        "orc_csr_copy": (None, [P, IP, IP, DP]),
        "orc_csr_from_arrays": (P, [I, I, IP, IP, DP, I]),
        "orc_csr_multiply": (I, [P, DP, DP]),
        "orc_csr_expanded_symmetric": (P, [P]),
        "orc_csr_lower_triangle": (P, [P]),
        "orc_assemble_solids": (P, [I, I, IP, IP, IP, DP, IP, DP, DP, I]),
        "orc_pcg_jacobi": (I, [P, DP, D, I, I, DP, IP, DP]),
        "orc_pcg_jacobi_fixed": (I, [P, DP, I, I, DP, DP]),
        "orc_true_relative_residual": (D, [P, DP, DP]),
        "orc_pload4_hexa8": (I, [DP, I, I, DP, DP]),
        "orc_quad4_ke": (I, [DP, D, D, D, D, D, D, I, DP]),
        "orc_quad4_me": (I, [DP, D, D, DP]),
        "orc_assemble_model": (P, [I, I, IP, IP, IP, DP, IP, DP, DP, DP, I, I]),
        "orc_quad4_dke_dt": (I, [DP, D, D, D, D, D, D, I, DP]),
        "orc_element_sensitivity": (I, [I, I, IP, IP, IP, DP, IP, DP, DP, DP, I, I, I, DP, DP, DP]),
        "orc_grav_rhs": (I, [I, I, IP, IP, IP, DP, IP, DP, DP, DP, DP]),
        "orc_quad4_recover": (I, [DP, D, D, D, D, I, DP, D, D, DP]),
        "orc_element_forces": (I, [I, IP, IP, IP, DP, IP, DP, DP, DP, I, I, DP, I, DP]),
        "orc_recover_stress": (I, [I, IP, IP, IP, DP, IP, DP, DP, DP, I, DP, DP, DP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _err():
    return lib().orc_last_error().decode()


# ---------------------------------------------------------------- DofMap
class DofMap:
    """Restatement of vibestran::DofMap (src/core/dof_map.cpp)."""

    def __init__(self, node_ids, dofs_per_node=6):
        ids = _i32(node_ids)
        self._h = lib().orc_dofmap_build(len(ids), _ip(ids), dofs_per_node)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_dofmap_free(self._h)
            self._h = None

    def constrain(self, node, dof):
        if lib().orc_dofmap_constrain(self._h, node, dof) != 0:
            raise OracleError(_err())

    def constrain_batch(self, pairs):
        n = _i32([p[0] for p in pairs])
        d = _i32([p[1] for p in pairs])
        if lib().orc_dofmap_constrain_batch(self._h, len(n), _ip(n), _ip(d)) != 0:
            raise OracleError(_err())

    def eq_index(self, node, dof):
        return lib().orc_dofmap_eq_index(self._h, node, dof)

    def is_free(self, node, dof):
        return self.eq_index(node, dof) != -1

    def num_free_dofs(self):
        return lib().orc_dofmap_num_free(self._h)

    def num_total_dofs(self):
        return lib().orc_dofmap_num_total(self._h)


def analysis_dof_map(node_ids, elem_type, elem_ptr, elem_nodes, spc_nodes, spc_masks):
    """build_analysis_dof_map restatement. Returns (eq[nnodes,6], n_free)."""
    ids = _i32(node_ids)
    et, ep, en = _i32(elem_type), _i32(elem_ptr), _i32(elem_nodes)
    sn, sm = _i32(spc_nodes), _i32(spc_masks)
    eq = np.empty((len(ids), 6), dtype=np.int32)
    nf = lib().orc_analysis_dof_map(len(ids), _ip(ids), len(et), _ip(et), _ip(ep), _ip(en),
                                    len(sn), _ip(sn), _ip(sm), _ip(eq))
    if nf < 0:
        raise OracleError(_err())
    return eq, nf


# ---------------------------------------------------------------- elements
def _elem(fn, xyz, *args, nd):
    x = _f64(xyz).reshape(-1)
    out = np.empty(nd * nd, dtype=np.float64)
    if fn(_dp(x), *args, _dp(out)) != 0:
        raise OracleError(_err())
    return out.reshape(nd, nd)


def hexa8_ke(xyz, E, nu, formulation="EAS"):
    f = lib().orc_hexa8_ke_sri if formulation == "SRI" else lib().orc_hexa8_ke_eas
    return _elem(f, xyz, E, nu, nd=24)


def tetra4_ke(xyz, E, nu):
    return _elem(lib().orc_tetra4_ke, xyz, E, nu, nd=12)


def penta6_ke(xyz, E, nu):
    return _elem(lib().orc_penta6_ke_eas, xyz, E, nu, nd=18)


def hexa8_me(xyz, rho):
    return _elem(lib().orc_hexa8_me, xyz, rho, nd=24)


def tetra4_me(xyz, rho):
    return _elem(lib().orc_tetra4_me, xyz, rho, nd=12)


def penta6_me(xyz, rho):
    return _elem(lib().orc_penta6_me, xyz, rho, nd=18)


def quad4_ke(xyz, E, nu, t, formulation="MITC4", G=0.0, twelveI_t3=1.0, tst=0.833333):
    return _elem(lib().orc_quad4_ke, xyz, E, nu, G, t, twelveI_t3, tst,
                 1 if formulation == "MINDLIN" else 0, nd=24)


def quad4_dke_dt(xyz, E, nu, t, formulation="MITC4", G=0.0, twelveI_t3=1.0, tst=0.833333):
    return _elem(lib().orc_quad4_dke_dt, xyz, E, nu, G, t, twelveI_t3, tst,
                 1 if formulation == "MINDLIN" else 0, nd=24)


def quad4_me(xyz, rho, t):
    return _elem(lib().orc_quad4_me, xyz, rho, t, nd=24)


def pload4_hexa8(xyz, g1_local=-1, g34_local=-1, pressures=(1.0, 1.0, 1.0, 1.0)):
    x = _f64(xyz).reshape(-1)
    p = _f64(pressures)
    out = np.empty(24, dtype=np.float64)
    if lib().orc_pload4_hexa8(_dp(x), g1_local, g34_local, _dp(p), _dp(out)) != 0:
        raise OracleError(_err())
    return out.reshape(8, 3)


# ---------------------------------------------------------------- CSR
class Csr:
    """Restatement of SparseMatrixBuilder::CsrData."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_csr_free(self._h)
            self._h = None

    @classmethod
    def from_arrays(cls, n, row_ptr, col_ind, values, lower_only):
        rp, ci, v = _i32(row_ptr), _i32(col_ind), _f64(values)
        return cls(lib().orc_csr_from_arrays(n, len(ci), _ip(rp), _ip(ci), _dp(v), int(lower_only)))

    @property
    def n(self):
        return lib().orc_csr_n(self._h)

    @property
    def nnz(self):
        return lib().orc_csr_nnz(self._h)

    def stores_lower_triangle_only(self):
        return bool(lib().orc_csr_is_lower(self._h))

    def arrays(self):
        rp = np.empty(self.n + 1, dtype=np.int32)
        ci = np.empty(self.nnz, dtype=np.int32)
        v = np.empty(self.nnz, dtype=np.float64)
        lib().orc_csr_copy(self._h, _ip(rp), _ip(ci), _dp(v))
        return rp, ci, v

    @property
    def row_ptr(self):
        return self.arrays()[0]

    @property
    def col_ind(self):
        return self.arrays()[1]

    @property
    def values(self):
        return self.arrays()[2]

    def multiply(self, x):
        x = _f64(x)
        y = np.empty(self.n, dtype=np.float64)
        lib().orc_csr_multiply(self._h, _dp(x), _dp(y))
        return y

    def expanded_symmetric(self):
        return Csr(lib().orc_csr_expanded_symmetric(self._h))

    def lower_triangle(self):
        return Csr(lib().orc_csr_lower_triangle(self._h))

    def to_scipy(self):
        import scipy.sparse as sp
        rp, ci, v = self.expanded_symmetric().arrays()
        return sp.csr_matrix((v, ci, rp), shape=(self.n, self.n))


class SparseMatrixBuilder:
    def __init__(self, n):
        self._h = lib().orc_builder_create(n)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_builder_free(self._h)
            self._h = None

    def add(self, row, col, value):
        lib().orc_builder_add(self._h, row, col, float(value))

    def add_element_stiffness(self, gdofs, ke):
        g = _i32(gdofs)
        k = _f64(ke).reshape(-1)
        if len(k) != len(g) * len(g):
            raise ValueError("ke size mismatch")
        lib().orc_builder_add_element(self._h, len(g), _ip(g), _dp(k))

    def num_triplets(self):
        return lib().orc_builder_num_triplets(self._h)

    def build_csr(self):
        return Csr(lib().orc_builder_build_csr(self._h))


def assemble_solids(n, elem_type, elem_ptr, conn, xyz, eq, E, nu, formulation="EAS"):
    et, ep, cn = _i32(elem_type), _i32(elem_ptr), _i32(conn)
    x, q = _f64(xyz).reshape(-1), _i32(eq).reshape(-1)
    e, v = _f64(E), _f64(nu)
    h = lib().orc_assemble_solids(n, len(et), _ip(et), _ip(ep), _ip(cn), _dp(x), _ip(q), _dp(e),
                                  _dp(v), 1 if formulation == "SRI" else 0)
    if not h:
        raise OracleError(_err())
    return Csr(h)


def assemble_model(n, model, eq):
    """Assemble a vibestran_amd Model (solids and CQUAD4) through the oracle."""
    et, ep, cn = _i32(model.elem_type), _i32(model.elem_ptr), _i32(model.conn)
    x, q = _f64(model.xyz).reshape(-1), _i32(eq).reshape(-1)
    e, v = _f64(model.E), _f64(model.nu)
    th = _f64(model.thickness if model.thickness is not None else np.zeros(len(et)))
    h = lib().orc_assemble_model(n, len(et), _ip(et), _ip(ep), _ip(cn), _dp(x), _ip(q), _dp(e),
                                 _dp(v), _dp(th), int(model.solid_form), int(model.shell_form))
    if not h:
        raise OracleError(_err())
    return Csr(h)


def grav_rhs(n, model, eq, g, rho=None):
    """F = sum_e M_e a (consistent masses), a = g at every node."""
    et, ep, cn = _i32(model.elem_type), _i32(model.elem_ptr), _i32(model.conn)
    x, q = _f64(model.xyz).reshape(-1), _i32(eq).reshape(-1)
    r = _f64(model.rho if rho is None else rho)
    th = _f64(model.thickness if model.thickness is not None else np.zeros(len(et)))
    F = np.empty(n, dtype=np.float64)
    if lib().orc_grav_rhs(n, len(et), _ip(et), _ip(ep), _ip(cn), _dp(x), _ip(q), _dp(r), _dp(th),
                          _dp(_f64(g)), _dp(F)) != 0:
        raise OracleError(_err())
    return F


def element_sensitivity(n, model, eq, param, u, v):
    """out[e] = -u_e^T (dK_e/dp_e) v_e with p = E (param 0) or CQUAD4
    thickness (param 1); the FP64 CPU adjoint contraction."""
    et, ep, cn = _i32(model.elem_type), _i32(model.elem_ptr), _i32(model.conn)
    x, q = _f64(model.xyz).reshape(-1), _i32(eq).reshape(-1)
    e, nv = _f64(model.E), _f64(model.nu)
    th = _f64(model.thickness if model.thickness is not None else np.zeros(len(et)))
    out = np.empty(len(et), dtype=np.float64)
    rc = lib().orc_element_sensitivity(n, len(et), _ip(et), _ip(ep), _ip(cn), _dp(x), _ip(q),
                                       _dp(e), _dp(nv), _dp(th), int(model.solid_form),
                                       int(model.shell_form), int(param), _dp(_f64(u)),
                                       _dp(_f64(v)), _dp(out))
    if rc != 0:
        raise OracleError(_err())
    return out


def quad4_recover(xyz, E, nu, t, ue, xi, eta, formulation="MITC4", twelveI_t3=1.0):
    """recover_shell_response at (xi, eta): (sx, sy, sxy, mx, my, mxy)."""
    x = _f64(xyz).reshape(-1)
    u = _f64(ue).reshape(-1)
    out = np.empty(6, dtype=np.float64)
    rc = lib().orc_quad4_recover(_dp(x), E, nu, t, twelveI_t3,
                                 0 if formulation == "MITC4" else 1, _dp(u), xi, eta, _dp(out))
    if rc != 0:
        raise OracleError(_err())
    return out


def element_forces(model, eq, u):
    """sum_e K_e u_e at every node DOF, [nn, 6]."""
    et, ep, cn = _i32(model.elem_type), _i32(model.elem_ptr), _i32(model.conn)
    x, q = _f64(model.xyz).reshape(-1), _i32(eq).reshape(-1)
    e, nv = _f64(model.E), _f64(model.nu)
    th = _f64(model.thickness if model.thickness is not None else np.zeros(len(et)))
    f = np.empty((model.nn, 6), dtype=np.float64)
    rc = lib().orc_element_forces(len(et), _ip(et), _ip(ep), _ip(cn), _dp(x), _ip(q), _dp(e),
                                  _dp(nv), _dp(th), int(model.solid_form), int(model.shell_form),
                                  _dp(_f64(u)), model.nn, _dp(f))
    if rc != 0:
        raise OracleError(_err())
    return f


STRESS_COLUMNS = ("sx", "sy", "sxy", "sz", "syz", "szx", "mx", "my", "mxy", "von_mises")


def recover_stress(model, eq, u, corners=True):
    """Element stresses of recover_results: (centroid [nelem, 10], corner
    [nelem, 8, 10] or None), columns STRESS_COLUMNS."""
    et, ep, cn = _i32(model.elem_type), _i32(model.elem_ptr), _i32(model.conn)
    x, q = _f64(model.xyz).reshape(-1), _i32(eq).reshape(-1)
    e, nv = _f64(model.E), _f64(model.nu)
    th = _f64(model.thickness if model.thickness is not None else np.zeros(len(et)))
    cen = np.empty((len(et), 10), dtype=np.float64)
    cor = np.empty((len(et), 8, 10), dtype=np.float64) if corners else None
    rc = lib().orc_recover_stress(len(et), _ip(et), _ip(ep), _ip(cn), _dp(x), _ip(q), _dp(e),
                                  _dp(nv), _dp(th), int(model.shell_form), _dp(_f64(u)),
                                  _dp(cen), _dp(cor) if corners else None)
    if rc != 0:
        raise OracleError(_err())
    return cen, cor


# ---------------------------------------------------------------- PCG
def pcg_jacobi(K: Csr, F, tol=1e-8, max_iters=10000, nthreads=0):
    """Returns (u, iters, rel_res); raises OracleError like SolverError."""
    F = _f64(F)
    u = np.empty(K.n, dtype=np.float64)
    it = C.c_int(0)
    rel = C.c_double(0.0)
    rc = lib().orc_pcg_jacobi(K._h, _dp(F), tol, max_iters, nthreads, _dp(u), C.byref(it),
                              C.byref(rel))
    if rc != 0:
        raise OracleError(_err())
    return u, it.value, rel.value


def pcg_jacobi_fixed(K: Csr, F, iters, nthreads=0):
    F = _f64(F)
    u = np.empty(K.n, dtype=np.float64)
    sec = C.c_double(0.0)
    rc = lib().orc_pcg_jacobi_fixed(K._h, _dp(F), iters, nthreads, _dp(u), C.byref(sec))
    if rc != 0:
        raise OracleError(_err())
    return u, sec.value


def true_relative_residual(K: Csr, u, F):
    return lib().orc_true_relative_residual(K._h, _dp(_f64(u)), _dp(_f64(F)))
