"""Oracle CQUAD4 pinned by the reference's element tests (symmetry, PSD,
rigid-body modes, orientation invariance: tests/unit/test_elements.cpp:85-150,
921-1100, 1539-1875) and closed-form membrane/bending states."""
import numpy as np
import pytest

import oracle as O

PLATE = np.array([[0, 0, 0], [2.0, 0, 0], [2.2, 1.1, 0], [-0.1, 0.9, 0]], float)


def rot(ax, ang):
    c, s = np.cos(ang), np.sin(ang)
    if ax == "x":
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    if ax == "y":
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def rigid6(X, axes=np.eye(3)):
    """Rigid translations and rotations about the rows of `axes`."""
    R = np.zeros((24, 6))
    for a, p in enumerate(X):
        R[6 * a:6 * a + 3, 0:3] = np.eye(3)
        for j, w in enumerate(axes):
            R[6 * a:6 * a + 3, 3 + j] = np.cross(w, p)
            R[6 * a + 3:6 * a + 6, 3 + j] = w
    return R


def shell_axes(X):
    n = np.cross(X[1] - X[0], X[3] - X[0])
    n /= np.linalg.norm(n)
    e1 = (X[1] - X[0]) / np.linalg.norm(X[1] - X[0])
    return np.array([e1, np.cross(n, e1), n])


@pytest.mark.parametrize("form", ["MITC4", "MINDLIN"])
@pytest.mark.parametrize("Q", [np.eye(3), rot("x", 0.7) @ rot("z", 0.3), rot("y", 1.2)])
def test_shell_invariants(form, Q):
    X = PLATE @ Q.T
    K = O.quad4_ke(X, 7e4, 0.33, 0.5, form)
    s = np.abs(K).max()
    assert np.abs(K - K.T).max() < 1e-10 * s
    KR = np.abs(K @ rigid6(X, shell_axes(X))).max(axis=0) / s
    # translations and the two in-plane rotations are exact zero-energy modes;
    # the rotation about the normal only meets the drilling penalty
    # (1e-6 x the largest rotational diagonal, cquad4.cpp:109-129)
    assert KR[:5].max() < 1e-12
    assert 1e-9 < KR[5] < 1e-6
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert ev.min() > -1e-12 * s and np.sum(ev > 1e-9 * s) == 18
    # spectrum is orientation independent (test_elements.cpp:1681-1709)
    K0 = O.quad4_ke(PLATE, 7e4, 0.33, 0.5, form)
    assert np.allclose(np.linalg.eigvalsh(0.5 * (K0 + K0.T)), ev, rtol=1e-9, atol=1e-9 * s)


def test_membrane_uniform_stretch_energy():  # test_elements.cpp:1936-1957
    X = np.array([[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 1, 0]], float)
    E, nu, t = 1e6, 0.3, 0.1
    for form in ("MITC4", "MINDLIN"):
        K = O.quad4_ke(X, E, nu, t, form)
        u = np.zeros(24)
        u[0::6] = 1e-3 * X[:, 0]  # exx = 1e-3, plane stress with free lateral: use exx only
        U = 0.5 * u @ K @ u
        # energy of strain (exx, 0, 0): 0.5 * t * A * E/(1-nu^2) * exx^2
        assert np.isclose(U, 0.5 * t * 2.0 * E / (1 - nu ** 2) * 1e-6, rtol=1e-12)


def test_pure_bending_energy():
    X = np.array([[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 1, 0]], float)
    E, nu, t = 1e6, 0.3, 0.1
    K = O.quad4_ke(X, E, nu, t, "MITC4")
    kappa = 1e-3
    u = np.zeros(24)
    for a, (x, y, z) in enumerate(X):
        u[6 * a + 2] = -0.5 * kappa * x * x   # w = -k x^2 / 2
        u[6 * a + 4] = kappa * x              # ry = -dw/dx
    U = 0.5 * u @ K @ u
    D = E * t ** 3 / (12 * (1 - nu ** 2))
    assert np.isclose(U, 0.5 * D * kappa ** 2 * 2.0, rtol=1e-6)


def test_drilling_penalty_is_tiny_and_positive():
    K = O.quad4_ke(PLATE, 7e4, 0.33, 0.5, "MITC4")
    rot = [6 * a + 3 for a in range(4)]
    drill = [6 * a + 5 for a in range(4)]
    assert all(K[d, d] > 0 for d in drill)
    assert max(K[d, d] for d in drill) < 1e-3 * max(K[r, r] for r in rot)
