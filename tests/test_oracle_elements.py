"""Oracle element matrices pinned by invariants (symmetry, rigid-body null
space, PSD: reference tests/unit/test_elements.cpp:328-920), by the constant
strain patch test on distorted meshes (exact for SRI and for EAS, whose modes
integrate to zero), and by the e2e closed form of chexa8_axial_force
(delta = 1e-3 exactly, Poisson -3e-4)."""
import json
import os

import numpy as np
import pytest

import oracle as O
from vibestran_amd import bdf, decks
from vibestran_amd import model as M

GOLD = os.path.join(os.path.dirname(__file__), "golden", "e2e")

CUBE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
WEDGE = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]], float)
TET = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)


def distorted(X, seed, amp=0.12):
    rng = np.random.default_rng(seed)
    return X + rng.uniform(-amp, amp, X.shape)


def rigid_modes(X):
    n = len(X)
    R = np.zeros((3 * n, 6))
    for a, (x, y, z) in enumerate(X):
        R[3 * a:3 * a + 3, 0:3] = np.eye(3)
        R[3 * a:3 * a + 3, 3] = [0, -z, y]
        R[3 * a:3 * a + 3, 4] = [z, 0, -x]
        R[3 * a:3 * a + 3, 5] = [-y, x, 0]
    return R


KINDS = [
    ("hexa-eas", lambda X, E, nu: O.hexa8_ke(X, E, nu, "EAS"), CUBE),
    ("hexa-sri", lambda X, E, nu: O.hexa8_ke(X, E, nu, "SRI"), CUBE),
    ("penta-eas", O.penta6_ke, WEDGE),
    ("tetra", O.tetra4_ke, TET),
]


@pytest.mark.parametrize("name,fn,X0", KINDS)
@pytest.mark.parametrize("seed", [0, 1])
def test_invariants(name, fn, X0, seed):
    X = X0 if seed == 0 else distorted(X0, seed)
    K = fn(X, 2.1e5, 0.3)
    s = np.abs(K).max()
    assert np.abs(K - K.T).max() < 1e-10 * s
    assert np.abs(K @ rigid_modes(X)).max() < 1e-9 * s
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert ev.min() > -1e-9 * s
    nz = len(ev) - 6
    assert np.sum(ev > 1e-8 * s) == nz  # exactly six zero-energy modes
    K2 = fn(X, 4.2e5, 0.3)
    assert np.allclose(K2, 2 * K, rtol=1e-13, atol=1e-12 * s)


def test_eas_softer_than_sri_near_incompressible():  # test_elements.cpp:873-919, 643-672
    Ke = O.hexa8_ke(CUBE, 1e6, 0.4999, "EAS")
    Ks = O.hexa8_ke(CUBE, 1e6, 0.4999, "SRI")
    # bending-type mode: antisymmetric x displacement through z
    u = np.zeros(24)
    for a, (x, y, z) in enumerate(CUBE):
        u[3 * a] = (z - 0.5)
    assert u @ Ke @ u <= u @ Ks @ u * (1 + 1e-12)
    assert np.linalg.eigvalsh(Ke).min() >= -1e-6 * np.abs(Ke).max()


def test_tetra_volume_closed_form():
    K = O.tetra4_ke(TET, 1.0, 0.0)  # nu=0: D = diag(1,1,1,.5,.5,.5)
    # uniform x-stretch u = x: strain exx = 1, energy = 0.5*E*V = 1/12
    u = np.zeros(12)
    u[0::3] = TET[:, 0]
    assert np.isclose(0.5 * u @ K @ u, 0.5 * 1.0 * (1 / 6), rtol=1e-14)


@pytest.mark.parametrize("form,mesh", [("EAS", "jitter"), ("EAS", "affine"), ("SRI", "affine")])
def test_hexa_patch_test_distorted_mesh(form, mesh):
    # EAS passes on arbitrary hexahedra (its modes integrate to zero exactly);
    # the SRI one-point volumetric rule is exact only for parallelepipeds.
    m = decks.hex_block(3, 3, 3, jitter=0.15 if mesh == "jitter" else 0.0, jitter_seed=5)
    if mesh == "affine":
        m.xyz = m.xyz @ np.array([[1.0, 0.3, 0.1], [0.0, 0.9, -0.2], [0.15, 0.0, 1.2]]).T
    m.solid_form = 1 if form == "SRI" else 0
    m.spc_nodes = np.zeros(0, np.int32)
    m.spc_masks = np.zeros(0, np.int32)
    eq, n = M.analysis_dof_map(m)
    K = O.assemble_solids(n, m.elem_type, m.elem_ptr, m.conn, m.xyz, eq, m.E, m.nu, form)
    A = np.array([[1e-3, 2e-4, -1e-4], [3e-4, -5e-4, 2e-4], [0.0, 1e-4, 4e-4]])
    U = m.xyz @ A.T
    u = np.zeros(n)
    u[eq[:, :3].reshape(-1)] = U.reshape(-1)
    f = K.multiply(u).reshape(-1, 3)
    ijk = m.meta["ijk"]
    interior = np.all((ijk > 0) & (ijk < 3), axis=1)
    scale = np.abs(f).max()
    assert np.abs(f[interior]).max() < 1e-10 * scale
    assert np.abs(f.sum(axis=0)).max() < 1e-10 * scale  # self-equilibrated


def test_penta_and_tetra_patch_test():
    # two wedges and five tets filling unit cubes, constant strain
    wedges = [WEDGE, np.array([[1, 0, 0], [1, 1, 0], [0, 1, 0], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)]
    A = np.array([[1e-3, 0, 2e-4], [0, -4e-4, 0], [1e-4, 0, 3e-4]])
    f = np.zeros((8, 3))
    nodes = {tuple(p): i for i, p in enumerate(CUBE)}
    for W in wedges:
        K = O.penta6_ke(distorted(W, 9, 0.0), 1e5, 0.25)
        u = (W @ A.T).reshape(-1)
        fe = (K @ u).reshape(-1, 3)
        for a, p in enumerate(W):
            f[nodes[tuple(p)]] += fe[a]
    # equilibrium of the assembled pair
    assert np.abs(f.sum(axis=0)).max() < 1e-10 * np.abs(f).max()


def test_mass_matrices_total_mass():  # test_elements.cpp:1403-1510
    X = distorted(CUBE, 4)
    Me = O.hexa8_me(X, 7.85e-9)
    V = np.linalg.det(np.array([X[1] - X[0], X[3] - X[0], X[4] - X[0]]))  # approx only
    tot = Me[0::3, 0::3].sum()
    assert tot > 0 and np.isclose(O.hexa8_me(CUBE, 2.0)[0::3, 0::3].sum(), 2.0, rtol=1e-14)
    assert np.isclose(O.tetra4_me(TET, 6.0)[0::3, 0::3].sum(), 1.0, rtol=1e-14)
    assert np.isclose(O.penta6_me(WEDGE, 2.0)[0::3, 0::3].sum(), 1.0, rtol=1e-14)
    assert V > 0


def _solve_deck(name):
    m = bdf.read_bdf(os.path.join(GOLD, name + ".bdf"))
    eq, n = M.analysis_dof_map(m)
    form = "SRI" if m.solid_form == 1 else "EAS"
    K = O.assemble_solids(n, m.elem_type, m.elem_ptr, m.node_ids[m.conn] * 0 + m.conn, m.xyz, eq,
                          m.E, m.nu, form)
    F = M.assemble_rhs(m, eq, n, m.loads[0])
    u, it, rel = O.pcg_jacobi(K, F, tol=1e-12)
    return m, eq, u


@pytest.mark.parametrize("deck", ["chexa8_axial_force", "ctetra4_axial_force"])
def test_e2e_solid_axial_decks(deck):
    m, eq, u = _solve_deck(deck)
    exp = json.load(open(os.path.join(GOLD, deck + ".expected.json")))
    idx = {int(g): i for i, g in enumerate(m.node_ids)}
    dofs = {"T1": 0, "T2": 1, "T3": 2}
    nchk = 0
    for c in exp["checks"]:
        if c["type"] != "node_displacement":
            continue
        e = eq[idx[c["node_id"]], dofs[c["dof"]]]
        val = 0.0 if e < 0 else u[e]
        if "rel_tol" in c:
            assert abs(val - c["expected"]) <= c["rel_tol"] * abs(c["expected"])
        else:
            assert abs(val - c["expected"]) <= c["abs_tol"]
        nchk += 1
    assert nchk >= 1


# ---- pins of the two EAS formulations ----------------------------------------------
# Pure bending. On a rectangular prism the trilinear field plus the nine
# enhanced modes contains the exact 3D pure-bending field (sigma_xx = E k z,
# every other stress 0):
#     u = k x z,   v = -nu k y z,   w = -k/2 (x^2 + nu (z^2 - y^2)),
# whose quadratic parts are exactly the incompatible bubbles the modes stand
# for (reference solid_elements.cpp:1095-1163: modes xi, eta, zeta in each of
# x, y, z). The exact stress is orthogonal to every enhanced mode, so the
# condensed Ke applied to the nodal values must return the consistent nodal
# forces of the end tractions +-sigma_xx on the x faces -- exactly. SRI (no
# enhancement), a mode set missing one direction, or a wrong condensation
# (Kuu - Kua Kaa^-1 Kau with a sign or transpose slip) leaves an O(1) residual.


def _bending_field(X, kap, nu):
    x, y, z = X.T
    return np.stack([kap * x * z, -nu * kap * y * z,
                     -0.5 * kap * (x ** 2 + nu * (z ** 2 - y ** 2))], axis=1).reshape(-1)


def _box_end_forces(X, a, b, c, E, kap):
    """Consistent nodal forces of sigma_xx = E k z on the faces x = +-a/2 of the
    box [-a/2, a/2] x [-b/2, b/2] x [-c/2, c/2] (2x2 Gauss on each face, exact for
    the bilinear x linear integrand)."""
    f = np.zeros((len(X), 3))
    g = np.array([-1.0, 1.0]) / np.sqrt(3.0)
    for sgn in (-1.0, 1.0):
        face = [i for i in range(len(X)) if np.isclose(X[i, 0], sgn * a / 2)]
        for s in g:
            for t in g:
                z = t * c / 2
                for i in face:
                    Ny = 0.5 * (1 + np.sign(X[i, 1]) * s)
                    Nz = 0.5 * (1 + np.sign(X[i, 2]) * t)
                    f[i, 0] += sgn * E * kap * z * Ny * Nz * (b / 2) * (c / 2)
    return f.reshape(-1)


@pytest.mark.parametrize("nu", [0.0, 0.3, 0.45])
def test_hexa_eas_pure_bending_exact(nu):
    E, kap, (a, b, c) = 2.1e5, 1e-3, (2.0, 1.5, 1.0)
    X = (2 * CUBE - 1) * np.array([a / 2, b / 2, c / 2])  # centred box, CHEXA8 order
    f = _box_end_forces(X, a, b, c, E, kap)
    u = _bending_field(X, kap, nu)
    r_eas = O.hexa8_ke(X, E, nu, "EAS") @ u - f
    assert np.abs(r_eas).max() < 1e-12 * np.abs(f).max()
    # the check has teeth: SRI (no enhanced modes) misses the bending field by O(1)
    r_sri = O.hexa8_ke(X, E, nu, "SRI") @ u - f
    assert np.abs(r_sri).max() > 0.2 * np.abs(f).max()


def test_penta_eas_pure_bending_exact():
    # Wedge of the unit right triangle x unit height, bending about y with
    # sigma_zz = E k x (nu = 0): u = -k z^2 / 2, w = k x z. The z^2 term is the
    # zeta bubble; the in-plane field is linear, so the EAS wedge (modes
    # L1 - 1/3, L2 - 1/3, zeta; solid_elements.cpp:1333-1467) is exact.
    E, kap = 7.0e4, 2e-3
    X = WEDGE.copy()
    x, z = X[:, 0], X[:, 2]
    u = np.stack([-0.5 * kap * z ** 2, 0 * x, kap * x * z], axis=1).reshape(-1)
    # consistent forces of +-sigma_zz on the two triangles (3-point midpoint
    # rule: exact for L_a * x, a quadratic)
    f = np.zeros((6, 3))
    mids = [(0.5, 0.5, 0.0), (0.0, 0.5, 0.5), (0.5, 0.0, 0.5)]  # (L_node0, L_node1, L_node2)
    tri = X[:3, :2]
    for top, sgn in ((0, -1.0), (3, 1.0)):
        for L in mids:
            xq = sum(L[k] * tri[k, 0] for k in range(3))
            for k in range(3):
                f[top + k, 2] += sgn * E * kap * xq * L[k] * 0.5 / 3
    f = f.reshape(-1)
    r = O.penta6_ke(X, E, 0.0) @ u - f
    assert np.abs(r).max() < 1e-12 * np.abs(f).max()


# Frame objectivity. A formulation whose enhanced-strain directions are the
# physical gradients of the parametric coordinates (the COLUMNS of J0^-1,
# d xi_k / dx) is rotation covariant: Ke(R X) = R_b Ke(X) R_b^T with R_b the
# block-diagonal node rotation. The reference's CHEXA8 EAS takes the ROWS of
# J0^-1 instead (solid_elements.cpp:1131-1133), which rotate as R J0^-1 only
# row-mixed: under a general rotation of a cube the mode space changes, while
# under a symmetric rotation (180 degrees about any axis, R = R^T) rows and
# columns coincide and covariance returns. CPENTA6 EAS uses the columns
# (1333-1467) and is covariant on any wedge. Swapping rows and columns in
# either element fails one of the assertions below.

def _rot_generic():
    w = np.array([0.3, -0.5, 0.7])
    th = np.linalg.norm(w)
    k = w / th
    Kx = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(th) * Kx + (1 - np.cos(th)) * Kx @ Kx


def _rot_half_turn():
    n = np.array([1.0, 2.0, -0.5])
    n /= np.linalg.norm(n)
    return 2 * np.outer(n, n) - np.eye(3)


def _covariance_defect(fn, X, R):
    K = fn(X)
    Kr = fn(X @ R.T)
    Rb = np.kron(np.eye(len(X)), R)
    return np.abs(Kr - Rb @ K @ Rb.T).max() / np.abs(K).max()


SHEAR = np.array([[1.0, 0.3, 0.2], [0.0, 1.0, 0.25], [0.0, 0.0, 1.0]])


def test_hexa_eas_uses_rows_of_inverse_centroidal_jacobian():
    hexa = lambda X: O.hexa8_ke(X, 2.1e5, 0.3, "EAS")  # noqa: E731
    cube = 2 * CUBE - 1
    assert _covariance_defect(hexa, cube, _rot_half_turn()) < 1e-12
    assert _covariance_defect(hexa, cube, _rot_generic()) > 1e-2
    assert _covariance_defect(hexa, cube @ SHEAR.T, _rot_generic()) > 1e-2
    # SRI has no enhanced modes: covariant on the same shapes
    sri = lambda X: O.hexa8_ke(X, 2.1e5, 0.3, "SRI")  # noqa: E731
    assert _covariance_defect(sri, cube @ SHEAR.T, _rot_generic()) < 1e-12


@pytest.mark.parametrize("shape", ["right", "sheared", "distorted"])
def test_penta_eas_is_rotation_covariant(shape):
    X = {"right": WEDGE, "sheared": WEDGE @ SHEAR.T,
         "distorted": distorted(WEDGE, 11, 0.08)}[shape]
    penta = lambda Y: O.penta6_ke(Y, 2.1e5, 0.3)  # noqa: E731
    assert _covariance_defect(penta, X, _rot_generic()) < 1e-12
    assert _covariance_defect(penta, X, _rot_half_turn()) < 1e-12
