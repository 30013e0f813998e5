"""Block PCG over k load cases through the C ABI: every column follows the
single-load-case recurrence, so each column is checked against the CPU oracle
solve of the same right-hand side (FP64: same iteration count within 2 and
solution within the PCG tolerance; mixed: FP64 true residual below 1e-8)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from vibestran_amd import CsrData, SolverError, _lib as L, decks
from vibestran_amd import model as M
from vibestran_amd.assembly import assemble_model, dof_map_device

pytestmark = pytest.mark.gpu


def csr_of(oc):
    rp, ci, v = oc.arrays()
    return CsrData.from_arrays(rp, ci, v, lower_only=oc.stores_lower_triangle_only())


@pytest.fixture(scope="module")
def c1():
    m = decks.config1_hex_cantilever()
    eq, n = M.analysis_dof_map(m)
    oc = O.assemble_model(n, m, eq)
    rng = np.random.default_rng(21)
    F = rng.standard_normal((n, 6))
    F[:, 0] = M.assemble_rhs(m, eq, n, m.loads[0])
    return m, eq, n, oc, F


@pytest.mark.parametrize("fmt", [-1, 0, 3])
def test_block_fp64_matches_single_solves(gpu_backend_factory, c1, fmt):
    m, eq, n, oc, F = c1
    be = gpu_backend_factory(tolerance=1e-10, max_iters=20000)
    be.set_format(fmt)
    K = csr_of(oc)
    U = be.solve_block(K, F)
    its = be.last_block_iterations
    for j in range(F.shape[1]):
        uo, it_o, _ = O.pcg_jacobi(oc, F[:, j], tol=1e-10, max_iters=20000)
        assert abs(int(its[j]) - it_o) <= 2, (j, its[j], it_o)
        assert np.abs(U[:, j] - uo).max() <= 1e-8 * np.abs(uo).max()
        assert O.true_relative_residual(oc, U[:, j], F[:, j]) < 1e-9
    cfg = be.block_config()
    assert cfg.ldk == 8 and cfg.col_tile in (4, 8) and cfg.spmm_ms > 0


def test_block_mixed_reaches_1e8(gpu_backend_factory, c1):
    m, eq, n, oc, F = c1
    be = gpu_backend_factory(mixed=True, tolerance=1e-8, max_iters=20000)
    U = be.solve_block(csr_of(oc), F)
    for j in range(F.shape[1]):
        rel = O.true_relative_residual(oc, U[:, j], F[:, j])
        assert rel < 1e-8 and abs(rel - be.last_block_residuals[j]) <= 1e-3 * rel + 1e-15
    assert be.stats().refinements >= F.shape[1]


def test_block_zero_and_repeated_columns(gpu_backend_factory, c1):
    m, eq, n, oc, F = c1
    be = gpu_backend_factory(tolerance=1e-10, max_iters=20000)
    G = np.stack([F[:, 1], np.zeros(n), F[:, 1], 2.0 * F[:, 1]], axis=1)
    U = be.solve_block(csr_of(oc), G)
    its = be.last_block_iterations
    assert its[1] == 0 and np.all(U[:, 1] == 0.0) and be.last_block_residuals[1] == 0.0
    assert np.array_equal(U[:, 0], U[:, 2]) and its[0] == its[2]
    assert np.abs(U[:, 3] - 2.0 * U[:, 0]).max() <= 1e-8 * np.abs(U[:, 3]).max()


@pytest.mark.parametrize("k", [1, 3, 32, 64])
def test_block_widths(gpu_backend_factory, k):
    oc = O.SparseMatrixBuilder(300)
    for i in range(300):
        oc.add(i, i, 4.0 + (i % 7))
        if i:
            oc.add(i, i - 1, -1.0)
        if i > 5:
            oc.add(i, i - 6, -0.5)
    A = oc.build_csr()
    be = gpu_backend_factory(tolerance=1e-12, max_iters=5000)
    F = np.random.default_rng(k).standard_normal((300, k))
    U = be.solve_block(csr_of(A), F)
    Ad = A.to_scipy().toarray()
    Ad = np.tril(Ad) + np.tril(Ad, -1).T
    X = np.linalg.solve(Ad, F)
    assert np.abs(U - X).max() <= 1e-9 * np.abs(X).max()


def test_block_errors(gpu_backend_factory, c1):
    m, eq, n, oc, F = c1
    be = gpu_backend_factory(max_iters=5)
    with pytest.raises(SolverError, match="load case 0: did not converge after 5 iterations"):
        be.solve_block(csr_of(oc), F[:, :2])
    assert np.all(be.last_block_iterations == 5)
    be2 = gpu_backend_factory()
    with pytest.raises(ValueError, match="k must be in 1..64"):
        be2.solve_block(csr_of(oc), np.zeros((n, 65)))


def test_block_device_pointers_tube(gpu_backend_factory):
    """CQUAD4 tube (BSR6), device-assembled operator, device-resident F/U."""
    be = gpu_backend_factory(tolerance=1e-9, max_iters=50000)
    m = decks.cquad4_tube(n_circ=24, n_axial=12, n_cases=8)
    eq, n = dof_map_device(be, m)
    K = assemble_model(be, m, eq, n)
    F = np.asfortranarray(np.stack([M.assemble_rhs(m, eq, n, lc) for lc in m.loads], axis=1))
    lib = L.load()
    nbytes = F.nbytes
    dF, dU = C.c_void_p(), C.c_void_p()
    assert lib.vbh_device_alloc(be._h, nbytes, C.byref(dF)) == 0
    assert lib.vbh_device_alloc(be._h, nbytes, C.byref(dU)) == 0
    try:
        assert lib.vbh_memcpy_h2d(be._h, dF, F.ctypes.data_as(C.c_void_p), nbytes) == 0
        its, rels = be.solve_block_device(dF.value, dU.value, F.shape[1])
        U = np.empty_like(F)
        assert lib.vbh_memcpy_d2h(be._h, U.ctypes.data_as(C.c_void_p), dU, nbytes) == 0
    finally:
        lib.vbh_device_free(be._h, dF)
        lib.vbh_device_free(be._h, dU)
    oc = O.assemble_model(n, m, eq)
    for j in range(F.shape[1]):
        uo, it_o, _ = O.pcg_jacobi(oc, F[:, j], tol=1e-9, max_iters=50000)
        # shells are ill-conditioned: the SpMM summation order moves the
        # count by a few iterations in a thousand
        assert abs(int(its[j]) - it_o) <= max(2, it_o // 50), (j, its[j], it_o)
        assert np.abs(U[:, j] - uo).max() <= 1e-6 * np.abs(uo).max()
        assert O.true_relative_residual(oc, U[:, j], F[:, j]) < 2e-9


def _variants():
    out = []
    for cw, lg in [(2, 16), (4, 8), (8, 4), (1, 32)]:
        for xcd in (0, 1):
            # block values: 0 direct loads, 1 through LDS, 2 broadcast across
            # the group's lanes (groups of 4, 8 or 16 lanes)
            for stage in ((0, 1, 2) if lg in (4, 8, 16) else (0, 1)):
                out.append(f"{cw},{lg},1024,{xcd},{stage}")
    # 3: the sliced ELL, one lane per row, 8 or 4 columns per lane
    for kc in (8, 4, 2):
        for xcd in (0, 1):
            out.append(f"{kc},1,1024,{xcd},3")
    return out


@pytest.mark.parametrize("cfg", _variants())
def test_block_every_spmm_variant(gpu_backend_factory, monkeypatch, cfg):
    """Each SpMM shape the autotuner can pick, forced, on a BSR6 shell
    operator (FP64 and mixed) and a BSR3 solid operator (mixed), k = 32."""
    monkeypatch.setenv("VBH_BLOCK_TUNE", cfg)
    m = decks.cquad4_tube(n_circ=16, n_axial=24, n_cases=32)
    be = gpu_backend_factory(tolerance=1e-10, max_iters=50000)
    eq, n = dof_map_device(be, m)
    K = assemble_model(be, m, eq, n)
    F = np.asfortranarray(np.stack([M.assemble_rhs(m, eq, n, lc) for lc in m.loads], axis=1))
    U = be.solve_block(K, F)
    cfg_used = be.block_config()
    assert (cfg_used.cols_per_lane, cfg_used.lanes_per_group, cfg_used.value_path) == \
        tuple(int(v) for v in (cfg.split(",")[:2] + cfg.split(",")[4:]))
    oc = O.assemble_model(n, m, eq)
    for j in (0, 13, 31):
        assert O.true_relative_residual(oc, U[:, j], F[:, j]) < 2e-10
    bm = gpu_backend_factory(mixed=True, tolerance=1e-8, max_iters=50000)
    Um = bm.solve_block(csr_of(oc), F)
    for j in (0, 13, 31):
        assert O.true_relative_residual(oc, Um[:, j], F[:, j]) < 1e-8
    hm = decks.config2_hex_jitter(nx=10, ny=4, nz=4, seed=7)
    eq2, n2 = M.analysis_dof_map(hm)
    oc2 = O.assemble_model(n2, hm, eq2)
    G = np.random.default_rng(3).standard_normal((n2, 32))
    be2 = gpu_backend_factory(mixed=True, tolerance=1e-8, max_iters=50000)
    U2 = be2.solve_block(csr_of(oc2), G)
    for j in (0, 7, 31):
        assert O.true_relative_residual(oc2, U2[:, j], G[:, j]) < 1e-8


@pytest.mark.parametrize("k,cfg", [(8, "1,8,1024,0,2"), (8, "2,4,1024,1,2"), (16, "2,8,1024,0,2"),
                                   (16, "1,16,1024,1,2"), (64, "4,16,1024,0,2"),
                                   (64, "8,8,1024,1,2"), (5, "2,4,1024,0,2"),
                                   (8, "8,1,2048,1,3"), (5, "4,1,1024,0,3"),
                                   (16, "8,1,1024,0,3"), (64, "4,1,2048,1,3"),
                                   (2, "2,1,1024,0,3")])
def test_block_broadcast_widths(gpu_backend_factory, monkeypatch, k, cfg):
    """The register-broadcast SpMM at the other column widths (groups of 4,
    8 and 16 lanes; k = 5 pads to 8 columns), and the sliced-ELL SpMM (value
    path 3, 8 or 4 columns per lane): the shell tube (BSR6) in FP64
    and mixed precision and a jittered hex (BSR3) in mixed precision, each
    column to the tolerance against the FP64 oracle operator."""
    monkeypatch.setenv("VBH_BLOCK_TUNE", cfg)
    m = decks.cquad4_tube(n_circ=16, n_axial=24, n_cases=k)
    be = gpu_backend_factory(tolerance=1e-10, max_iters=50000)
    eq, n = dof_map_device(be, m)
    K = assemble_model(be, m, eq, n)
    F = np.asfortranarray(np.stack([M.assemble_rhs(m, eq, n, lc) for lc in m.loads], axis=1))
    U = be.solve_block(K, F)
    used = be.block_config()
    assert (used.cols_per_lane, used.lanes_per_group, used.value_path) == \
        tuple(int(v) for v in (cfg.split(",")[:2] + cfg.split(",")[4:]))
    oc = O.assemble_model(n, m, eq)
    for j in sorted({0, k // 2, k - 1}):
        assert O.true_relative_residual(oc, U[:, j], F[:, j]) < 2e-10
    bm = gpu_backend_factory(mixed=True, tolerance=1e-8, max_iters=50000)
    Um = bm.solve_block(csr_of(oc), F)
    for j in sorted({0, k // 2, k - 1}):
        assert O.true_relative_residual(oc, Um[:, j], F[:, j]) < 1e-8
    hm = decks.config2_hex_jitter(nx=10, ny=4, nz=4, seed=7)
    eq2, n2 = M.analysis_dof_map(hm)
    oc2 = O.assemble_model(n2, hm, eq2)
    G = np.random.default_rng(3).standard_normal((n2, k))
    U2 = gpu_backend_factory(mixed=True, tolerance=1e-8, max_iters=50000).solve_block(csr_of(oc2), G)
    for j in sorted({0, k // 2, k - 1}):
        assert O.true_relative_residual(oc2, U2[:, j], G[:, j]) < 1e-8


@pytest.mark.parametrize("mixed", [False, True])
def test_block_column_compaction(gpu_backend_factory, c1, monkeypatch, mixed):
    """32 load cases of which 20 are zero (done at once) and the rest converge
    at different iterations: the running columns move into 16-, then 8-wide
    layouts. Each column must match its own single-case solve and the
    full-width block run (VBH_BLOCK_COMPACT=0); zero columns stay zero."""
    m, eq, n, oc, F = c1
    rng = np.random.default_rng(5)
    G = np.zeros((n, 32))
    live = rng.choice(32, 12, replace=False)
    for i, j in enumerate(live):
        G[:, j] = F[:, i % 6] * (1.0 + i) if i < 6 else rng.standard_normal(n) * 10.0 ** (i - 8)
    tol = 1e-8 if mixed else 1e-10
    runs = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("VBH_BLOCK_COMPACT", flag)
        be = gpu_backend_factory(mixed=mixed, tolerance=tol, max_iters=20000)
        U = be.solve_block(csr_of(oc), G)
        runs[flag] = (U, np.array(be.last_block_iterations), np.array(be.last_block_residuals))
    (U, its, rels), (U0, its0, rels0) = runs["1"], runs["0"]
    dead = np.setdiff1d(np.arange(32), live)
    assert np.all(U[:, dead] == 0.0) and np.all(its[dead] == 0) and np.all(rels[dead] == 0.0)
    for j in live:
        rel = O.true_relative_residual(oc, U[:, j], G[:, j])
        assert rel < tol * (1.0 if mixed else 10.0), (j, rel)
        assert abs(int(its[j]) - int(its0[j])) <= max(3, 0.02 * its0[j]), (j, its[j], its0[j])
        scale = np.abs(U0[:, j]).max()
        assert np.abs(U[:, j] - U0[:, j]).max() <= (1e-6 if mixed else 1e-8) * scale, j
        if not mixed:
            uo, it_o, _ = O.pcg_jacobi(oc, G[:, j], tol=tol, max_iters=20000)
            assert abs(int(its[j]) - it_o) <= 2, (j, its[j], it_o)


@pytest.mark.parametrize("rel", ["1", "0"])
def test_block_sliced_ell_layouts(gpu_backend_factory, monkeypatch, rel):
    """The sliced-ELL SpMM reads whichever single-RHS layout the operator
    carries: the relative one (slot offsets per slice) or, with
    VBH_SELL_REL=0, the absolute one with 16-bit block columns, and with
    32-bit columns on a hex whose shuffled node IDs put neighbours more than
    32,767 nodes apart. FP64 and mixed block solves of 8 load cases, every
    checked column to the tolerance against the oracle operator."""
    monkeypatch.setenv("VBH_SELL_REL", rel)
    monkeypatch.setenv("VBH_BLOCK_TUNE", "8,1,1024,1,3")
    hx = decks.config2_hex_jitter(nx=64, ny=24, nz=24, seed=3)
    hx.node_ids = np.random.default_rng(9).permutation(hx.node_ids).astype(np.int32)
    cases = [decks.config2_hex_jitter(nx=10, ny=4, nz=4, seed=7),
             decks.cquad4_tube(n_circ=16, n_axial=24, n_cases=1), hx]
    for m in cases:
        eq, n = M.analysis_dof_map(m)
        oc = O.assemble_model(n, m, eq)
        G = np.random.default_rng(11).standard_normal((n, 8))
        for mixed, tol, bound in ((False, 1e-10, 2e-10), (True, 1e-8, 1e-8)):
            be = gpu_backend_factory(mixed=mixed, tolerance=tol, max_iters=50000)
            U = be.solve_block(csr_of(oc), G)
            used = be.block_config()
            assert (used.cols_per_lane, used.lanes_per_group, used.value_path) == (8, 1, 3)
            for j in (0, 3, 7):
                assert O.true_relative_residual(oc, U[:, j], G[:, j]) < bound
