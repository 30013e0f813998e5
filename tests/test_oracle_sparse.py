"""Oracle CSR builder pinned by the reference's own known answers
(tests/unit/test_sparse_matrix.cpp:49-178) and brute-force dense checks."""
import numpy as np

import oracle as O


def test_build_csr_diagonal():  # test_sparse_matrix.cpp:49-61
    b = O.SparseMatrixBuilder(2)
    b.add(0, 0, 3.0)
    b.add(1, 1, 5.0)
    c = b.build_csr()
    assert c.n == 2 and c.nnz == 2
    assert list(c.row_ptr) == [0, 1, 2]


def test_merge_and_duplicates_sum():  # test_sparse_matrix.cpp:63-90
    b = O.SparseMatrixBuilder(3)
    b.add(0, 0, 1.25)
    b.add(1, 2, -2.0)
    b.add(0, 0, 0.75)
    b.add(2, 1, 3.5)
    c = b.build_csr()
    rp, ci, v = c.arrays()
    assert c.nnz == 3
    assert list(rp) == [0, 1, 2, 3]
    assert list(ci) == [0, 2, 1]
    assert list(v) == [2.0, -2.0, 3.5]
    assert not c.stores_lower_triangle_only()


def test_element_assembly_lower_triangle():  # test_sparse_matrix.cpp:92-114
    b = O.SparseMatrixBuilder(2)
    b.add_element_stiffness([0, 1], [4.0, -1.5, -1.5, 3.0])
    c = b.build_csr()
    rp, ci, v = c.arrays()
    assert c.stores_lower_triangle_only() and c.nnz == 3
    assert list(rp) == [0, 1, 3] and list(ci) == [0, 0, 1]
    assert list(v) == [4.0, -1.5, 3.0]


def test_lower_matvec_matches_expanded():  # test_sparse_matrix.cpp:116-142
    b = O.SparseMatrixBuilder(2)
    b.add_element_stiffness([0, 1], [4.0, -1.5, -1.5, 3.0])
    low = b.build_csr()
    full = low.expanded_symmetric()
    yl = low.multiply([2.0, -1.0])
    yf = full.multiply([2.0, -1.0])
    assert full.nnz == 4 and not full.stores_lower_triangle_only()
    assert yl[0] == 9.5 and yl[1] == -6.0
    assert np.array_equal(yl, yf)


def test_lower_triangle_extraction():  # test_sparse_matrix.cpp:144-178
    b = O.SparseMatrixBuilder(3)
    for r, c, v in [(0, 0, 4.0), (0, 1, -1.0), (1, 0, -1.0), (1, 1, 3.0), (1, 2, -2.0),
                    (2, 1, -2.0), (2, 2, 5.0)]:
        b.add(r, c, v)
    full = b.build_csr()
    low = full.lower_triangle()
    rp, ci, v = low.arrays()
    assert low.stores_lower_triangle_only() and low.n == 3 and low.nnz == 5
    assert list(rp) == [0, 1, 3, 5]
    assert list(ci) == [0, 0, 1, 1, 2]
    assert list(v) == [4.0, -1.0, 3.0, -2.0, 5.0]


def test_exact_zero_entries_dropped_and_constrained_skipped():  # sparse_matrix.cpp:16-35
    b = O.SparseMatrixBuilder(3)
    ke = np.array([[2.0, 0.0, 1.0], [0.0, 3.0, 0.0], [1.0, 0.0, 4.0]])
    b.add_element_stiffness([2, 0, -1], ke)
    c = b.build_csr()
    rp, ci, v = c.arrays()
    # local (1,0) -> global (0,2) is upper: skipped; its mirror (2,0) has value 0.0: dropped
    assert list(rp) == [0, 1, 1, 2]
    assert list(ci) == [0, 2] and list(v) == [3.0, 2.0]


def test_random_assembly_matches_dense_bruteforce():
    rng = np.random.default_rng(0)
    n = 30
    b = O.SparseMatrixBuilder(n)
    K = np.zeros((n, n))
    for _ in range(40):
        g = rng.choice(np.arange(-1, n), size=6, replace=False)
        A = rng.standard_normal((6, 6))
        A[rng.random((6, 6)) < 0.2] = 0.0
        ke = A + A.T
        b.add_element_stiffness(g, ke)
        for i in range(6):
            for j in range(6):
                if g[i] >= 0 and g[j] >= 0 and g[i] >= g[j] and ke[i, j] != 0.0:
                    K[g[i], g[j]] += ke[i, j]
    c = b.build_csr()
    rp, ci, v = c.arrays()
    D = np.zeros((n, n))
    for r in range(n):
        for k in range(rp[r], rp[r + 1]):
            D[r, ci[k]] = v[k]
        assert np.all(np.diff(ci[rp[r]:rp[r + 1]]) > 0)
    assert np.allclose(D, K, rtol=0, atol=1e-12)
    x = rng.standard_normal(n)
    Kf = K + np.tril(K, -1).T
    assert np.allclose(c.multiply(x), Kf @ x, atol=1e-10)
