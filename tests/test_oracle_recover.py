"""Oracle stress recovery (recover_results element part,
linear_static.cpp:676-1197) pinned by closed forms and by the reference's e2e
fixtures:
  * constant strain (affine displacement) on distorted solids: every centroid
    and corner row is lam tr(eps) I + 2 mu eps;
  * fields the element interpolates exactly (trilinear on a box, bilinear x
    linear on a right prism): each corner row is the stress of the field's
    gradient at that node, which pins the corner order;
  * CQUAD4 membrane (affine in-plane field) and Kirchhoff bending
    (w = (kx x^2 + ky y^2 + kxy x y) / 2 with theta = (dw/dy, -dw/dx)):
    stress = Dp eps, moment = t^3/12 Dp kappa, unchanged by a rigid rotation
    of element and field (outputs live in the element frame);
  * von Mises from principal stresses (eigenvalues), not from the component
    formula;
  * the element_stress checks of chexa8_axial_force, cquad4_axial_force and
    ctetra4_axial_force, and every check of all e2e decks after a round trip
    through libvbh's CSV writer and the e2e runner's CSV reader.
"""
import os

import numpy as np
import pytest

import oracle as O
from e2e_csv import apply_checks, load_csv_results, uses_frames_or_mpc
from vibestran_amd import bdf
from vibestran_amd import model as M

GOLD = os.path.join(os.path.dirname(__file__), "golden", "e2e")
E, NU = 2.1e5, 0.3
LAM = E * NU / ((1 + NU) * (1 - 2 * NU))
MU = E / (2 * (1 + NU))


def one_element(et, X, t=1.0):
    X = np.asarray(X, float)
    nn = len(X)
    m = M.Model(node_ids=np.arange(1, nn + 1, dtype=np.int32), xyz=X,
                elem_type=np.array([et], np.int32), elem_ptr=np.array([0, nn], np.int32),
                conn=np.arange(nn, dtype=np.int32), E=np.array([E]), nu=np.array([NU]),
                rho=np.array([1.0]), thickness=np.array([t]) if et == M.CQUAD4 else None)
    eq = np.arange(6 * nn, dtype=np.int32).reshape(nn, 6)
    return m, eq


def solid_u(X, field):
    """6 DOFs per node (rotations 0) from a displacement field x -> u."""
    u = np.zeros((len(X), 6))
    for a, x in enumerate(X):
        u[a, :3] = field(x)
    return u.reshape(-1)


def solid_expected(H):
    """Row [sx, sy, sxy, sz, syz, szx, 0, 0, 0, vm] of the gradient H."""
    eps = 0.5 * (H + H.T)
    S = LAM * np.trace(eps) * np.eye(3) + 2 * MU * eps
    p = np.linalg.eigvalsh(S)
    vm = np.sqrt(0.5 * ((p[0] - p[1]) ** 2 + (p[1] - p[2]) ** 2 + (p[2] - p[0]) ** 2))
    return np.array([S[0, 0], S[1, 1], S[0, 1], S[2, 2], S[1, 2], S[0, 2], 0, 0, 0, vm])


CUBE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
WEDGE = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]], float)
TET = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
NCORNER = {M.CHEXA8: 8, M.CPENTA6: 6, M.CTETRA4: 4}


@pytest.mark.parametrize("et,X0", [(M.CHEXA8, CUBE), (M.CPENTA6, WEDGE), (M.CTETRA4, TET)])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_solid_constant_strain(et, X0, seed):
    rng = np.random.default_rng(seed)
    X = X0 * rng.uniform(0.5, 2.0) + (rng.uniform(-0.12, 0.12, X0.shape) if seed else 0.0)
    A = rng.normal(size=(3, 3)) * 1e-3
    c = rng.normal(size=3)
    m, eq = one_element(et, X)
    cen, cor = O.recover_stress(m, eq, solid_u(X, lambda x: A @ x + c))
    ref = solid_expected(A)
    scale = np.abs(ref).max()
    assert np.abs(cen[0] - ref).max() < 1e-9 * scale
    nc = NCORNER[et]
    for k in range(nc):
        assert np.abs(cor[0, k] - ref).max() < 1e-9 * scale, k
    assert np.all(cor[0, nc:] == 0.0)


def test_hexa_trilinear_field_corners():
    # box [0,2]x[0,3]x[0,1.5]: the trilinear space holds 1, x, y, z, xy, yz, zx, xyz
    X = CUBE * np.array([2.0, 3.0, 1.5])
    a, b, c, d = 1e-3, -2e-3, 1.5e-3, 0.7e-3

    def field(x):
        return np.array([a * x[0] * x[1] * x[2], b * x[1] * x[2] + d * x[0], c * x[2] * x[0]])

    def grad(x):
        return np.array([[a * x[1] * x[2], a * x[0] * x[2], a * x[0] * x[1]],
                         [d, b * x[2], b * x[1]],
                         [c * x[2], 0.0, c * x[0]]])

    m, eq = one_element(M.CHEXA8, X)
    cen, cor = O.recover_stress(m, eq, solid_u(X, field))
    scale = np.abs(solid_expected(grad(X[6]))).max()
    assert np.abs(cen[0] - solid_expected(grad(X.mean(axis=0)))).max() < 1e-10 * scale
    for k in range(8):
        assert np.abs(cor[0, k] - solid_expected(grad(X[k]))).max() < 1e-10 * scale, k


def test_penta_prism_field_corners():
    # right prism over the triangle (0,0),(2,0),(0.5,1.5), height 2: space {1,x,y} x {1,z}
    X = np.array([[0, 0, 0], [2, 0, 0], [0.5, 1.5, 0], [0, 0, 2], [2, 0, 2], [0.5, 1.5, 2]], float)

    def field(x):
        return np.array([1e-3 * x[0] * x[2], -2e-3 * x[1] * x[2] + 1e-3 * x[0], 5e-4 * x[2]])

    def grad(x):
        return np.array([[1e-3 * x[2], 0.0, 1e-3 * x[0]],
                         [1e-3, -2e-3 * x[2], -2e-3 * x[1]],
                         [0.0, 0.0, 5e-4]])

    m, eq = one_element(M.CPENTA6, X)
    cen, cor = O.recover_stress(m, eq, solid_u(X, field))
    scale = np.abs(solid_expected(grad(X[4]))).max()
    centroid = np.array([2.5 / 3, 0.5, 1.0])  # L1 = L2 = 1/3, zeta = 0
    assert np.abs(cen[0] - solid_expected(grad(centroid))).max() < 1e-10 * scale
    for k in range(6):
        assert np.abs(cor[0, k] - solid_expected(grad(X[k]))).max() < 1e-10 * scale, k


def rot(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def plate_expected(sig, mom):
    sx, sy, sxy = sig
    p = np.linalg.eigvalsh(np.array([[sx, sxy], [sxy, sy]]))
    vm = np.sqrt(p[0] ** 2 - p[0] * p[1] + p[1] ** 2)
    return np.array([sx, sy, sxy, 0, 0, 0, mom[0], mom[1], mom[2], vm])


# node 1 at the origin and node 2 on +x, so the element frame is the basic frame
QUADS = [np.array([[0, 0, 0], [2, 0, 0], [2, 1.5, 0], [0, 1.5, 0]], float),
         np.array([[0, 0, 0], [2.2, 0, 0], [1.9, 1.7, 0], [-0.2, 1.3, 0]], float),
         np.array([[0, 0, 0], [1.0, 0, 0], [1.4, 0.8, 0], [0.1, 1.1, 0]], float)]


def plate_dp(t):
    c = E / (1 - NU ** 2)
    Dp = np.array([[c, c * NU, 0], [c * NU, c, 0], [0, 0, c * (1 - NU) / 2]])
    return Dp, t ** 3 / 12 * Dp


@pytest.mark.parametrize("form", ["MITC4", "MINDLIN"])
@pytest.mark.parametrize("qi", range(3))
def test_quad4_membrane_and_bending(form, qi):
    X = QUADS[qi]
    t = 0.3
    m, eq = one_element(M.CQUAD4, X, t)
    m.shell_form = 0 if form == "MITC4" else 1
    Dp, Db = plate_dp(t)
    A = np.array([[1e-3, 4e-4], [-3e-4, -6e-4]])
    kx, ky, kxy = 2e-3, -1e-3, 3e-3
    u = np.zeros((4, 6))
    for a, x in enumerate(X):
        u[a, :2] = A @ x[:2] + [0.2, -0.1]
        u[a, 2] = 0.5 * (kx * x[0] ** 2 + ky * x[1] ** 2 + kxy * x[0] * x[1])  # w
        u[a, 3] = ky * x[1] + 0.5 * kxy * x[0]        # theta_x = dw/dy
        u[a, 4] = -(kx * x[0] + 0.5 * kxy * x[1])     # theta_y = -dw/dx
    eps = np.array([A[0, 0], A[1, 1], A[0, 1] + A[1, 0]])
    ref = plate_expected(Dp @ eps, Db @ np.array([kx, ky, kxy]))
    cen, cor = O.recover_stress(m, eq, u.reshape(-1))
    tol_s = 1e-9 * np.abs(ref[:3]).max()
    tol_m = 1e-9 * np.abs(ref[6:9]).max()
    for row in [cen[0]] + [cor[0, k] for k in range(4)]:
        assert np.abs(row[:3] - ref[:3]).max() < tol_s
        assert np.abs(row[6:9] - ref[6:9]).max() < tol_m
        assert np.all(row[3:6] == 0.0)
        assert abs(row[9] - ref[9]) < tol_s
    assert np.all(cor[0, 4:] == 0.0)
    # rigid rotation of element and field: element-frame outputs unchanged
    Q = rot(qi + 11)
    m2, _ = one_element(M.CQUAD4, X @ Q.T + [3.0, -1.0, 2.0], t)
    m2.shell_form = m.shell_form
    u2 = np.concatenate([u[:, :3] @ Q.T, u[:, 3:] @ Q.T], axis=1)
    cen2, cor2 = O.recover_stress(m2, eq, u2.reshape(-1))
    assert np.abs(cen2 - cen).max() < 1e-9 * np.abs(cen).max()
    assert np.abs(cor2 - cor).max() < 1e-9 * np.abs(cor).max()


def test_quad4_bilinear_membrane_field_on_rectangle():
    # u = a x y on a rectangle: sx = Dp00 a y and gamma_xy = a x at every point,
    # exact for the standard B and for MITC4+ (whose tied shear carries the
    # r (T_rr - B_rr) / 4 in-plane bending term)
    X = QUADS[0]
    a = 1e-3
    u = np.zeros((4, 6))
    u[:, 0] = a * X[:, 0] * X[:, 1]
    Dp, _ = plate_dp(0.3)
    for form in (0, 1):
        m, eq = one_element(M.CQUAD4, X, 0.3)
        m.shell_form = form
        cen, cor = O.recover_stress(m, eq, u.reshape(-1))
        pts = [X.mean(axis=0)] + [X[k] for k in range(4)]
        rows = [cen[0]] + [cor[0, k] for k in range(4)]
        for x, row in zip(pts, rows):
            s = np.array([Dp[0, 0] * a * x[1], Dp[1, 0] * a * x[1], Dp[2, 2] * a * x[0]])
            assert np.abs(row[:3] - s).max() < 1e-9 * Dp[0, 0] * a * 2, (form, x)


def _solve_and_recover(name):
    m = bdf.read_bdf(os.path.join(GOLD, name + ".bdf"))
    eq, n = M.analysis_dof_map(m)
    K = O.assemble_model(n, m, eq)
    F = M.assemble_rhs(m, eq, n, m.loads[0])
    u, _, _ = O.pcg_jacobi(K, F, tol=1e-12, max_iters=200000)
    cen, _ = O.recover_stress(m, eq, u, corners=False)
    disp = np.where(eq >= 0, u[np.maximum(eq, 0)], 0.0)
    return m, disp, cen


DECKS = sorted(f[:-4] for f in os.listdir(GOLD) if f.endswith(".bdf") and not uses_frames_or_mpc(f[:-4]))


@pytest.mark.parametrize("name", DECKS)
def test_e2e_decks_through_csv(name, tmp_path):
    from vibestran_amd import results as R
    m, disp, cen = _solve_and_recover(name)
    node_csv, elem_csv = R.write_csv(str(tmp_path / (name + ".op2")), m,
                                     [(m.subcase_id(0), disp, cen)])
    assert node_csv.endswith(name + ".node.csv") and elem_csv.endswith(name + ".elem.csv")
    node_data, elem_data = load_csv_results(node_csv, elem_csv)
    assert apply_checks(name, node_data, elem_data) > 0


def test_csv_format(tmp_path):
    from vibestran_amd import results as R
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1]], float)
    m = M.Model(node_ids=np.array([30, 10, 20, 50, 40], np.int32), xyz=X,
                elem_type=np.array([M.CTETRA4, M.CQUAD4], np.int32),
                elem_ptr=np.array([0, 4, 8], np.int32),
                conn=np.array([0, 1, 3, 4, 0, 1, 2, 3], np.int32), E=np.ones(2), nu=np.zeros(2),
                rho=np.ones(2), thickness=np.array([0.0, 1.0]),
                elem_ids=np.array([7, 3], np.int32))
    disp = np.arange(30, dtype=float).reshape(5, 6) * -1.25e-3
    rows = np.arange(20, dtype=float).reshape(2, 10) * 3.5e4
    rows[0, 2] = -0.0
    nc, ec = R.write_csv(str(tmp_path / "out"), m, [(1, disp, rows), (4, disp * 2, None)])
    nl = open(nc).read().splitlines()
    el = open(ec).read().splitlines()
    assert nl[0] == "# node_id, subcase_id, T1, T2, T3, R1, R2, R3"
    assert el[0] == "# elem_id, subcase_id, sx, sy, sxy, sz, syz, szx, mx, my, mxy, von_mises"
    order = np.argsort(m.node_ids)
    want = [f"{m.node_ids[i]}, 1, " + ", ".join("%.8e" % v for v in disp[i]) for i in order]
    want += [f"{m.node_ids[i]}, 4, " + ", ".join("%.8e" % v for v in 2 * disp[i]) for i in order]
    assert nl[1:] == want
    # plates first, then solids
    assert el[1:] == ["3, 1, " + ", ".join("%.8e" % v for v in rows[1]),
                      "7, 1, " + ", ".join("%.8e" % v for v in rows[0])]
    assert "-0.00000000e+00" in el[2]
    assert nl[3].startswith("30, 1, -0.00000000e+00, -1.25000000e-03")
