"""Pressure loads pinned by closed forms: the consistent nodal forces of a
PLOAD4 on any CHEXA8 face sum to -p times the outward area vector of the
(bilinear) face, 0.5 (d13 x d24); on a flat rectangle every corner carries
p A / 4; and the host RHS assembly equals the oracle's element forces."""
import numpy as np
import pytest

import oracle as O
from vibestran_amd import decks
from vibestran_amd import model as M

FACES = [[0, 1, 2, 3], [4, 5, 6, 7], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7]]


def jittered_hex(seed):
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float) * [2.0, 1.5, 1.0]
    return X + np.random.default_rng(seed).uniform(-0.1, 0.1, X.shape)


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("face", range(6))
def test_pload4_resultant_is_pressure_times_outward_area(seed, face):
    X = jittered_hex(seed)
    f = FACES[face]
    p = 3.5
    Fe = O.pload4_hexa8(X, f[0], f[2], (p, p, p, p))
    a = 0.5 * np.cross(X[f[2]] - X[f[0]], X[f[3]] - X[f[1]])
    if a @ (X[f].mean(axis=0) - X.mean(axis=0)) < 0:
        a = -a  # outward
    assert np.abs(Fe.sum(axis=0) - (-p) * a).max() <= 1e-12 * p * np.linalg.norm(a)
    others = [k for k in range(8) if k not in f]
    assert np.all(Fe[others] == 0.0)


def test_pload4_flat_rectangle_equal_corner_shares():
    X = np.array([[0, 0, 0], [2, 0, 0], [2, 3, 0], [0, 3, 0],
                  [0, 0, 1], [2, 0, 1], [2, 3, 1], [0, 3, 1]], float)
    Fe = O.pload4_hexa8(X, 4, 6, (2.0, 2.0, 2.0, 2.0))  # top face z = 1, outward +z
    for k in (4, 5, 6, 7):
        assert np.allclose(Fe[k], [0.0, 0.0, -2.0 * 6.0 / 4.0], atol=1e-13)


def test_host_rhs_matches_oracle_element_forces():
    m = decks.config2_hex_jitter(nx=5, ny=3, nz=3, seed=8)
    eq, n = M.analysis_dof_map(m)
    F = M.assemble_rhs(m, eq, n, m.loads[0])
    G = np.zeros(n)
    pl = m.loads[0].pload4
    c = m.conn2d()
    faces = M.hexa_pload4_faces(m, pl)
    for i, e in enumerate(pl.elem):
        Fe = O.pload4_hexa8(m.xyz[c[e]], int(faces[i][0]), int(faces[i][2]), pl.p[i])
        for a in range(8):
            for d in range(3):
                q = eq[c[e, a], d]
                if q >= 0:
                    G[q] += Fe[a, d]
    assert np.abs(F - G).max() <= 1e-12 * np.abs(G).max()
    top = np.abs(F).sum()
    assert top > 0
