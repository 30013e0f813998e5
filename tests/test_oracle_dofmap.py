"""Oracle DofMap pinned by tests/unit/test_dof_map.cpp:21-300 known answers,
and the product's closed-form numbering checked bit-exactly against it."""
import numpy as np

import oracle as O
from vibestran_amd import decks
from vibestran_amd import model as M


def test_two_nodes_twelve_dofs():  # test_dof_map.cpp:21-27
    dm = O.DofMap([1, 2])
    assert dm.num_free_dofs() == 12 and dm.num_total_dofs() == 12
    seen = {dm.eq_index(n, d) for n in (1, 2) for d in range(6)}
    assert seen == set(range(12))


def test_constrain_shifts_higher():  # test_dof_map.cpp:45-74
    dm = O.DofMap([1, 2])
    before = dm.eq_index(2, 0)
    dm.constrain(1, 0)
    assert dm.num_free_dofs() == 11 and dm.eq_index(1, 0) == -1
    assert dm.eq_index(2, 0) == before - 1


def test_unknown_node_is_constrained():  # test_dof_map.cpp:87-92
    dm = O.DofMap([1, 2])
    assert dm.eq_index(999, 0) == -1 and not dm.is_free(999, 0)


def test_constrain_batch_matches_serial():  # test_dof_map.cpp:160-191
    pairs = [(1, 3), (1, 4), (1, 5), (2, 3), (2, 4), (2, 5), (3, 0), (4, 2)]
    s = O.DofMap([1, 2, 3, 4])
    for n, d in pairs:
        s.constrain(n, d)
    b = O.DofMap([1, 2, 3, 4])
    b.constrain_batch(pairs)
    assert b.num_free_dofs() == s.num_free_dofs() == 16
    for n in range(1, 5):
        for d in range(6):
            assert b.eq_index(n, d) == s.eq_index(n, d)


def test_constrain_batch_ignores_duplicates():  # test_dof_map.cpp:194-208
    dm = O.DofMap([1, 2])
    dm.constrain_batch([(1, 0), (1, 0), (1, 1)])
    assert dm.num_free_dofs() == 10
    assert dm.eq_index(1, 0) == -1 and dm.eq_index(1, 1) == -1 and dm.eq_index(1, 2) != -1


def _analysis(m):
    return O.analysis_dof_map(m.node_ids, m.elem_type, m.elem_ptr, m.node_ids[m.conn],
                              m.node_ids[m.spc_nodes],
                              m.spc_masks)


def test_shell_solid_shared_node_keeps_rotations():  # test_dof_map.cpp:242-262
    ids = np.arange(1, 9, dtype=np.int32)
    et = np.array([O.CQUAD4, O.CTETRA4], np.int32)
    ep = np.array([0, 4, 8], np.int32)
    conn = np.array([1, 2, 4, 3, 1, 5, 6, 7], np.int32)
    eq, nf = O.analysis_dof_map(ids, et, ep, conn, [], [])
    assert (eq[0, 3:] >= 0).all()      # node 1 shared: rotations free
    assert eq[4, 3] == -1              # node 5 solid only: R1 inactive
    assert nf == 4 * 6 + 3 * 3         # nodes 1-4 shell (6), 5-7 solid (3), 8 unused


def test_closed_form_numbering_matches_constrain_batch():
    m = decks.hex_block(5, 3, 4)
    # shuffle node ids so sorted-id order differs from storage order
    rng = np.random.default_rng(3)
    m.node_ids = rng.permutation(np.arange(1, m.nn + 1) * 7).astype(np.int32)
    # partial SPCs on a few more nodes
    extra = rng.choice(m.nn, size=10, replace=False).astype(np.int32)
    m.spc_nodes = np.concatenate([m.spc_nodes, extra])
    m.spc_masks = np.concatenate([m.spc_masks, rng.integers(1, 64, size=10).astype(np.int32)])
    eq_o, nf_o = O.analysis_dof_map(m.node_ids, m.elem_type, m.elem_ptr, m.node_ids[m.conn],
                                    m.node_ids[m.spc_nodes], m.spc_masks)
    eq_p, nf_p = M.analysis_dof_map(m)
    assert nf_o == nf_p
    assert np.array_equal(eq_o, eq_p)
