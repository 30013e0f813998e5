"""Python mirror of the reference's solver plugin API over libvbh.

``HipPcgSolverBackend`` follows vibestran::SolverBackend
(reference include/solver/solver_backend.hpp:20-51) and the CUDA PCG
backend's factory/diagnostics (include/solver/cuda_pcg_solver_backend.hpp:29-76):
``try_create`` returns None when no GPU is usable, ``solve(K, F)`` returns a
new displacement vector and raises ``SolverError`` on every failure the
reference throws for.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib as L


class SolverError(RuntimeError):
    """vibestran::SolverError (reference include/core/exceptions.hpp:14-17)."""


@dataclass
class CsrData:
    """SparseMatrixBuilder::CsrData (reference include/core/sparse_matrix.hpp:63-87)."""

    row_ptr: np.ndarray
    col_ind: np.ndarray
    values: np.ndarray
    n: int
    nnz: int
    symmetry: str = "Full"  # "Full" or "Lower"
    _key: int = field(default=0, repr=False)

    def stores_lower_triangle_only(self) -> bool:
        return self.symmetry == "Lower"

    @classmethod
    def from_arrays(cls, row_ptr, col_ind, values, lower_only=True):
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(col_ind, dtype=np.int32)
        v = np.ascontiguousarray(values, dtype=np.float64)
        return cls(rp, ci, v, len(rp) - 1, len(ci), "Lower" if lower_only else "Full")


def _check(rc: int, what: str = ""):
    if rc == L.VBH_OK:
        return
    msg = L.last_error()
    if rc == L.VBH_ERR_SOLVER:
        raise SolverError(msg)
    if rc == L.VBH_ERR_ARG:
        raise ValueError(msg)
    raise RuntimeError(f"{what}: {msg}" if what else msg)


class HipPcgSolverBackend:
    """Jacobi-PCG on one MI355X through libvbh (``--backend=hip-pcg``)."""

    PRECISION = L.VBH_PREC_FP64

    def __init__(self, handle, tol, max_iters):
        self._h = handle
        self._tol = tol
        self._max_iters = max_iters
        self._uploaded = None
        self._last_iters = 0
        self._last_rel = 0.0
        self._lib = L.load()
        _check(self._lib.vbh_set_precision(self._h, self.PRECISION))

    # -- factory ------------------------------------------------------------
    @classmethod
    def try_create(cls, tolerance: float = 0.0, max_iters: int = 0,
                   device: int = 0) -> Optional["HipPcgSolverBackend"]:
        lib = L.load()
        h = C.c_void_p()
        tol = tolerance if tolerance > 0.0 else cls.default_tolerance()
        rc = lib.vbh_ctx_create(device, tol, max_iters, C.byref(h))
        if rc == L.VBH_ERR_NO_DEVICE:
            return None
        _check(rc, "vbh_ctx_create")
        return cls(h, tol, max_iters if max_iters > 0 else 10000)

    @staticmethod
    def default_tolerance() -> float:
        return 1e-8

    def close(self):
        if getattr(self, "_h", None):
            self._lib.vbh_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- SolverBackend ------------------------------------------------------
    def name(self) -> str:
        return "HIP PCG + Jacobi float64 (GPU)"

    def requires_full_symmetric_csr(self) -> bool:
        return False

    def last_iteration_count(self) -> int:
        return self._last_iters

    def last_estimated_error(self) -> float:
        return self._last_rel

    def last_relative_residual(self) -> float:
        return self._last_rel

    def device_name(self) -> str:
        return self._lib.vbh_device_name(self._h).decode()

    def set_format(self, fmt: int):
        """-1 auto, 0 scalar CSR, 3 / 6 block CSR."""
        _check(self._lib.vbh_set_format(self._h, fmt))
        self._uploaded = None

    def set_matrix(self, K: CsrData):
        if self._uploaded is K:
            return
        if hasattr(K, "fetch") and getattr(K, "_backend", None) is not self:
            K.fetch()  # device matrix of another context: go through the host arrays
        rp = np.ascontiguousarray(K.row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(K.col_ind, dtype=np.int32)
        v = np.ascontiguousarray(K.values, dtype=np.float64)
        _check(self._lib.vbh_set_csr(self._h, int(K.n), int(len(ci)), L.ip(rp), L.ip(ci), L.dp(v),
                                     1 if K.stores_lower_triangle_only() else 0))
        self._uploaded = K

    def solve(self, K: CsrData, F, factor_ratio_policy=None) -> np.ndarray:
        del factor_ratio_policy  # MAXRATIO applies to direct factorisations only
        F = np.ascontiguousarray(F, dtype=np.float64)
        if K.n == 0:
            raise SolverError("HIP PCG: stiffness matrix is empty -- no free DOFs")
        if F.shape[0] != K.n:
            raise SolverError(f"HIP PCG: force vector size {F.shape[0]} != matrix size {K.n}")
        self.set_matrix(K)
        u = np.empty(K.n, dtype=np.float64)
        it = C.c_int(0)
        rel = C.c_double(0.0)
        rc = self._lib.vbh_solve(self._h, L.dp(F), L.dp(u), C.byref(it), C.byref(rel))
        self._last_iters = it.value
        self._last_rel = rel.value
        _check(rc)
        return u

    def solve_block(self, K: CsrData, F) -> np.ndarray:
        """Solve K U = F for the columns of F (n x k) in one fused block-PCG."""
        F = np.asfortranarray(F, dtype=np.float64)
        if F.ndim != 2 or F.shape[0] != K.n:
            raise SolverError("HIP PCG: block force array must be n x k")
        k = F.shape[1]
        self.set_matrix(K)
        U = np.empty((K.n, k), dtype=np.float64, order="F")
        its = np.zeros(k, dtype=np.int32)
        rels = np.zeros(k, dtype=np.float64)
        rc = self._lib.vbh_solve_block(self._h, k, L.dp(F), L.dp(U), L.ip(its), L.dp(rels))
        self._last_iters = int(its.max()) if k else 0
        self._last_rel = float(rels.max()) if k else 0.0
        self.last_block_iterations = its
        self.last_block_residuals = rels
        _check(rc)
        return U

    def solve_block_device(self, dF: int, dU: int, k: int):
        """Block solve on device-resident column-major n x k arrays (HBM
        addresses); returns (iterations[k], rel_res[k])."""
        its = np.zeros(k, dtype=np.int32)
        rels = np.zeros(k, dtype=np.float64)
        rc = self._lib.vbh_solve_block_device(self._h, k, C.c_void_p(dF), C.c_void_p(dU),
                                              L.ip(its), L.dp(rels))
        self.last_block_iterations = its
        self.last_block_residuals = rels
        _check(rc)
        return its, rels

    def block_config(self) -> L.vbh_block_config:
        cfg = L.vbh_block_config()
        _check(self._lib.vbh_last_block_config(self._h, C.byref(cfg)))
        return cfg

    def spmv(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        _check(self._lib.vbh_spmv(self._h, L.dp(x), L.dp(y)))
        return y

    def stats(self) -> L.vbh_solve_stats:
        s = L.vbh_solve_stats()
        _check(self._lib.vbh_last_stats(self._h, C.byref(s)))
        return s


class HipMixedPcgSolverBackend(HipPcgSolverBackend):
    """FP32 SpMV/Jacobi inner iterations with FP64 true-residual refreshes and
    FP64 solution update (replaces cuda-pcg-mixed,
    reference cuda_pcg_backend_impl.hpp:2134-2262, without its 3-solve cap)."""

    PRECISION = L.VBH_PREC_MIXED

    @staticmethod
    def default_tolerance() -> float:
        return 1e-8

    def name(self) -> str:
        return "HIP PCG + Jacobi float32 / float64 refinement (GPU)"
