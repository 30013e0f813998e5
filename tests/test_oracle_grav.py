"""GRAV load vectors (consistent mass times a uniform acceleration) pinned by
closed forms: each node of an undistorted element carries its share of
rho V g (1/8 cube, 1/4 tetrahedron, 1/6 right prism, 1/4 rectangle shell), and
the resultant over a mesh is the total mass times g."""
import numpy as np

import oracle as O
from vibestran_amd import decks
from vibestran_amd import model as M

G = np.array([1.5, -2.0, -9810.0])


def one_element(et, X, rho=2.0, t=0.0):
    nn = len(X)
    m = M.Model(node_ids=np.arange(1, nn + 1, dtype=np.int32), xyz=np.asarray(X, float),
                elem_type=np.array([et], np.int32), elem_ptr=np.array([0, nn], np.int32),
                conn=np.arange(nn, dtype=np.int32), E=np.array([1e5]), nu=np.array([0.3]),
                rho=np.array([rho]), thickness=np.array([t]) if et == M.CQUAD4 else None,
                spc_nodes=np.zeros(0, np.int32), spc_masks=np.zeros(0, np.int32))
    eq = np.arange(6 * nn, dtype=np.int32).reshape(nn, 6)
    return O.grav_rhs(6 * nn, m, eq, G).reshape(nn, 6)


def test_single_element_shares():
    cube = [[0, 0, 0], [2, 0, 0], [2, 2, 0], [0, 2, 0], [0, 0, 2], [2, 0, 2], [2, 2, 2], [0, 2, 2]]
    f = one_element(M.CHEXA8, cube)
    assert np.allclose(f[:, :3], 2.0 * 8.0 / 8 * G, rtol=1e-14, atol=0)
    tet = [[0, 0, 0], [3, 0, 0], [0, 2, 0], [0, 0, 1]]
    f = one_element(M.CTETRA4, tet)
    assert np.allclose(f[:, :3], 2.0 * 1.0 / 4 * G, rtol=1e-14, atol=0)
    wedge = [[0, 0, 0], [2, 0, 0], [0, 1, 0], [0, 0, 3], [2, 0, 3], [0, 1, 3]]
    f = one_element(M.CPENTA6, wedge)
    assert np.allclose(f[:, :3], 2.0 * 3.0 / 6 * G, rtol=1e-14, atol=0)
    quad = [[0, 0, 0], [2, 0, 0], [2, 3, 0], [0, 3, 0]]
    f = one_element(M.CQUAD4, quad, t=0.5)
    assert np.allclose(f[:, :3], 2.0 * 0.5 * 6.0 / 4 * G, rtol=1e-14, atol=0)
    assert np.all(f[:, 3:] == 0.0)


def test_mesh_resultant_is_total_mass_times_g():
    hm = decks.hex_block(4, 2, 2)
    c = hm.conn2d()
    et, conn, vol = [], [], []
    for e in range(len(c)):
        if e % 3 == 0:
            et.append(M.CHEXA8); conn.extend(c[e]); vol.append(1.0)
        elif e % 3 == 1:
            for w in ([0, 1, 2, 4, 5, 6], [0, 2, 3, 4, 6, 7]):
                et.append(M.CPENTA6); conn.extend(c[e, w]); vol.append(0.5)
        else:
            for t in ([0, 1, 3, 4], [1, 2, 3, 6], [1, 4, 5, 6], [3, 4, 6, 7], [1, 3, 4, 6]):
                et.append(M.CTETRA4); conn.extend(c[e, t]); vol.append(1 / 3 if t == [1, 3, 4, 6] else 1 / 6)
    ep = np.cumsum([0] + [M.NODES_PER[t] for t in et]).astype(np.int32)
    rho = np.random.default_rng(0).uniform(1.0, 3.0, len(et))
    m = M.Model(node_ids=hm.node_ids, xyz=hm.xyz, elem_type=np.array(et, np.int32), elem_ptr=ep,
                conn=np.array(conn, np.int32), E=np.full(len(et), 1e5), nu=np.full(len(et), 0.3),
                rho=rho, spc_nodes=np.zeros(0, np.int32), spc_masks=np.zeros(0, np.int32))
    eq = np.arange(6 * m.nn, dtype=np.int32).reshape(-1, 6)
    f = O.grav_rhs(6 * m.nn, m, eq, G).reshape(-1, 6)
    assert np.allclose(f[:, :3].sum(axis=0), (rho * np.array(vol)).sum() * G, rtol=1e-13)
