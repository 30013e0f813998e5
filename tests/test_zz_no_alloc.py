"""The allocation-free solve (vbh_solve_device_noalloc) against the solve it
mirrors (vbh_solve_device), through the C ABI: two calls in a row, the second
on a smaller operator, in FP64 and in mixed precision. Both run the same
kernels in the same order on the same operator, so iteration counts and
solutions must be bit-identical; the solutions are also checked against the
oracle's operator (true relative residual below the tolerance). Operators
past the scratch array's size are rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from vibestran_amd import _lib, decks
from vibestran_amd import model as M
from vibestran_amd.assembly import assemble_model, dof_map_device

pytestmark = pytest.mark.gpu


def _dev(lib, h, a):
    p = C.c_void_p()
    assert lib.vbh_device_alloc(h, max(a.nbytes, 8), C.byref(p)) == 0
    assert lib.vbh_memcpy_h2d(h, p, a.ctypes.data_as(C.c_void_p), a.nbytes) == 0
    return p


def _solve(lib, h, fn, dF, n):
    du = C.c_void_p()
    assert lib.vbh_device_alloc(h, n * 8, C.byref(du)) == 0
    it, rel = C.c_int(-1), C.c_double(-1.0)
    rc = fn(h, dF, du, C.byref(it), C.byref(rel))
    assert rc == 0, _lib.last_error()
    u = np.empty(n)
    assert lib.vbh_memcpy_d2h(h, u.ctypes.data_as(C.c_void_p), du, u.nbytes) == 0
    lib.vbh_device_free(h, du)
    return u, it.value, rel.value


@pytest.mark.parametrize("precision", [_lib.VBH_PREC_FP64, _lib.VBH_PREC_MIXED])
def test_noalloc_solve_matches_solve_device(gpu_backend_factory, precision):
    lib = _lib.load()
    be = gpu_backend_factory(tolerance=1e-8, max_iters=50000)
    h = be._h
    assert lib.vbh_set_precision(h, precision) == 0
    # the second model is smaller than the first
    for m in (decks.config2_hex_jitter(nx=24, ny=8, nz=8, seed=5), decks.hex_block(5, 3, 2)):
        if not m.loads:
            tip = np.nonzero(m.meta["ijk"][:, 0] == m.meta["ijk"][:, 0].max())[0]
            nodal = np.zeros((m.nn, 6))
            nodal[tip, 2] = -10.0
            m.loads.append(M.LoadCase(nodal=nodal))
        eq, n = dof_map_device(be, m)
        assemble_model(be, m, eq, n)
        F = M.assemble_rhs(m, eq, n, m.loads[0])
        dF = _dev(lib, h, F)
        u_ref, it_ref, rel_ref = _solve(lib, h, lib.vbh_solve_device, dF, n)
        u, it, rel = _solve(lib, h, lib.vbh_solve_device_noalloc, dF, n)
        lib.vbh_device_free(h, dF)
        print(f"\n[noalloc] precision={precision} n={n}: iterations {it} (solve_device {it_ref}), "
              f"rel {rel:.3e}")
        assert it == it_ref and rel == rel_ref
        assert np.array_equal(u, u_ref)
        oc = O.assemble_solids(n, m.elem_type, m.elem_ptr, m.conn, m.xyz, eq, m.E, m.nu)
        assert O.true_relative_residual(oc, u, F) < 1e-8


def test_noalloc_rejects_operator_past_scratch(gpu_backend_factory):
    lib = _lib.load()
    be = gpu_backend_factory()
    h = be._h
    n = 2_600_000  # 2.6M rows x 48 bytes > 96 MiB
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    v = np.full(n, 2.0)
    assert lib.vbh_set_csr(h, n, n, rp.ctypes.data_as(C.POINTER(C.c_int)),
                           ci.ctypes.data_as(C.POINTER(C.c_int)),
                           v.ctypes.data_as(C.POINTER(C.c_double)), 1) == 0
    F = np.ones(n)
    dF = _dev(lib, h, F)
    du = C.c_void_p()
    assert lib.vbh_device_alloc(h, n * 8, C.byref(du)) == 0
    rc = lib.vbh_solve_device_noalloc(h, dF, du, None, None)
    msg = _lib.last_error()
    lib.vbh_device_free(h, dF)
    lib.vbh_device_free(h, du)
    assert rc == _lib.VBH_ERR_ARG and "scratch" in msg
