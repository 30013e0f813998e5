"""The headline solve, checked against the oracle: the mixed-precision
single-load-case Jacobi-PCG (FP32 SpMV / Jacobi, FP64 true-residual refresh
and FP64 solution) on the jittered, lognormal-E CHEXA8 block with PLOAD4 on
the top faces -- the 120x40x40 workload's generator at 40x16x16 (34,680 DOF),
a size the CPU oracle assembles and solves in about a second.

Checked: the device numbering and the device CSR against the oracle
(bit-identical pattern and values), the device PLOAD4 load vector
against the oracle's per-face pressure loads, the oracle's FP64 true relative
residual ||F - K u|| / ||F|| of the mixed solution (< 1e-8, the reference's
residual definition, linear_static.cpp:287-304), and the max-norm error of
the mixed solution against the oracle's FP64 Jacobi-PCG solution taken to
1e-12 (bounded by 1e-7 of max|u|), with the error of the plain FP64 GPU solve
reported beside it."""
import numpy as np
import pytest

import oracle as O
from vibestran_amd import CsrData, decks
from vibestran_amd import model as M
from vibestran_amd.assembly import assemble_model, assemble_rhs_device, dof_map_device

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2small():
    m = decks.config2_hex_jitter(nx=40, ny=16, nz=16)
    eq, n = M.analysis_dof_map(m)
    oc = O.assemble_solids(n, m.elem_type, m.elem_ptr, m.conn, m.xyz, eq, m.E, m.nu)
    F = np.zeros(n)
    pl = m.loads[0].pload4
    for e, g1, g34, p in zip(pl.elem, pl.g1, pl.g34, pl.p):
        nodes = m.conn[m.elem_ptr[e]:m.elem_ptr[e + 1]]
        fe = O.pload4_hexa8(m.xyz[nodes], int(g1), int(g34), p).reshape(8, 3)
        for a, node in enumerate(nodes):
            for d in range(3):
                if eq[node, d] >= 0:
                    F[eq[node, d]] += fe[a, d]
    u_ref, it_ref, _ = O.pcg_jacobi(oc, F, tol=1e-12, max_iters=100000)
    return m, eq, n, oc, F, u_ref, it_ref


def _max_err(u, u_ref):
    return float(np.abs(u - u_ref).max() / np.abs(u_ref).max())


def test_c2_workload_mixed_solve_vs_oracle(gpu_backend_factory, c2small):
    m, eq, n, oc, F, u_ref, it_ref = c2small
    be = gpu_backend_factory(mixed=True, tolerance=1e-8, max_iters=100000)
    eq_d, n_d = dof_map_device(be, m)
    assert n_d == n and np.array_equal(eq_d, eq)
    K = assemble_model(be, m, eq, n)
    Kh = K.fetch()
    rp, ci, v = oc.arrays()
    assert np.array_equal(Kh.row_ptr, rp) and np.array_equal(Kh.col_ind, ci)
    assert np.array_equal(Kh.values.view(np.int64), v.view(np.int64))  # bit-identical K
    Fd = assemble_rhs_device(be, m, eq, n, m.loads[0])
    assert np.abs(Fd - F).max() <= 1e-13 * np.abs(F).max()

    u = be.solve(K, Fd)
    rel = O.true_relative_residual(oc, u, F)
    err = _max_err(u, u_ref)
    st = be.stats()
    print(f"\n[c2 40x16x16 mixed] n={n} iters={st.iterations} refreshes={st.refinements} "
          f"oracle true rel={rel:.3e} max-norm error vs FP64 oracle={err:.3e} "
          f"(oracle FP64 PCG to 1e-12: {it_ref} iterations)")
    assert rel < 1e-8
    assert st.refinements >= 1
    assert err < 1e-7

    # plain FP64 GPU PCG on the same device operator, same bar
    be64 = gpu_backend_factory(mixed=False, tolerance=1e-8, max_iters=100000)
    K64 = CsrData.from_arrays(rp.copy(), ci.copy(), v.copy(), lower_only=True)
    u64 = be64.solve(K64, F)
    rel64 = O.true_relative_residual(oc, u64, F)
    err64 = _max_err(u64, u_ref)
    print(f"[c2 40x16x16 fp64 ] iters={be64.last_iteration_count()} oracle true rel={rel64:.3e} "
          f"max-norm error vs FP64 oracle={err64:.3e}; mixed vs fp64 GPU {_max_err(u, u64):.3e}")
    assert rel64 < 1e-8 and err64 < 1e-7


def test_c2_workload_slice_orders_vs_oracle(gpu_backend_factory, c2small, monkeypatch):
    """The FP32 SpMV of a 3x3-block operator sweeps XCD-contiguous slice
    ranges (and the vector kernel the matching row ranges) by default, the
    grid-stride order with VBH_SPMV_XCD=0. Both orders, mixed precision, meet
    the same oracle bars; they differ only in the summation order of the
    per-block partial sums, so their iteration counts stay within a few per
    cent of each other."""
    m, eq, n, oc, F, u_ref, it_ref = c2small
    rp, ci, v = oc.arrays()
    res = {}
    for xcd in ("1", "0"):
        monkeypatch.setenv("VBH_SPMV_XCD", xcd)
        be = gpu_backend_factory(mixed=True, tolerance=1e-8, max_iters=100000)
        u = be.solve(CsrData.from_arrays(rp.copy(), ci.copy(), v.copy(), lower_only=True), F)
        rel = O.true_relative_residual(oc, u, F)
        err = _max_err(u, u_ref)
        res[xcd] = be.last_iteration_count()
        print(f"\n[c2 40x16x16 mixed, VBH_SPMV_XCD={xcd}] iters={res[xcd]} oracle true rel={rel:.3e} "
              f"max-norm error vs FP64 oracle={err:.3e}")
        assert rel < 1e-8 and err < 1e-7
    assert abs(res["1"] - res["0"]) <= 0.1 * res["0"]
