"""Native BDF reader (libvbh vbh_deck_*, host only) against the reference's
BdfParser rules (src/io/bdf_parser.cpp): the e2e decks parse to the same model
as the pure-Python reader and their expected displacements follow through the
oracle; free, small and large field forms of one deck agree; Nastran number
forms, SPC1/PLOAD4 THRU lists, continuations, comments and case control give
the values the reference's rules fix; unsupported input fails with the
reference's "Line N: ..." messages. Parsing a deck large enough to be cut into
parallel chunks gives the same model as the sequential Python reader."""
import os

import numpy as np
import pytest

import oracle as O
from e2e_csv import GOLD, deck_checks, uses_frames_or_mpc
from vibestran_amd import bdf, deck, decks
from vibestran_amd import model as M
from vibestran_amd.backend import SolverError

DECKS = sorted(f[:-4] for f in os.listdir(GOLD) if f.endswith(".bdf"))
PY_DECKS = [d for d in DECKS if not uses_frames_or_mpc(d)]


def merged_spc(m):
    mask = np.zeros(m.nn, np.int64)
    np.bitwise_or.at(mask, m.spc_nodes, m.spc_masks)
    return mask


def same_model(a, b, exact=True):
    for k in ("node_ids", "elem_type", "elem_ptr", "conn"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    for k in ("xyz", "E", "nu", "rho"):
        if exact:
            assert np.array_equal(getattr(a, k), getattr(b, k)), k
        else:
            assert np.allclose(getattr(a, k), getattr(b, k), rtol=1e-9, atol=1e-12), k
    assert a.solid_form == b.solid_form and a.shell_form == b.shell_form
    assert np.array_equal(merged_spc(a), merged_spc(b))
    assert len(a.loads) == len(b.loads)
    for x, y in zip(a.loads, b.loads):
        assert np.array_equal(x.nodal, y.nodal)
        assert (x.pload4 is None) == (y.pload4 is None)
        if x.pload4 is not None:
            for k in ("elem", "g1", "g34", "p"):
                assert np.array_equal(np.asarray(getattr(x.pload4, k)),
                                      np.asarray(getattr(y.pload4, k))), k


@pytest.mark.parametrize("name", PY_DECKS)
def test_e2e_decks_match_python_reader(name):
    path = os.path.join(GOLD, name + ".bdf")
    a, b = bdf.read_bdf(path), deck.read_deck(path)
    same_model(a, b)
    if a.thickness is not None:
        assert np.array_equal(a.thickness, b.thickness)
    assert b.meta["unsupported_cards"] == []


@pytest.mark.parametrize("name", ["chexa8_axial_force", "ctetra4_axial_force",
                                  "cquad4_axial_force"])
def test_e2e_deck_displacements_through_native_reader(name):
    m = deck.read_deck(os.path.join(GOLD, name + ".bdf"))
    eq, n = M.analysis_dof_map(m)
    u, _, _ = O.pcg_jacobi(O.assemble_model(n, m, eq), M.assemble_rhs(m, eq, n, m.loads[0]),
                           tol=1e-12, max_iters=100000)
    idx = {int(g): i for i, g in enumerate(m.node_ids)}
    dofs = {"T1": 0, "T2": 1, "T3": 2, "R1": 3, "R2": 4, "R3": 5}
    checks = [c for c in deck_checks(name) if c["type"] == "node_displacement"]
    assert checks
    for c in checks:
        e = eq[idx[c["node_id"]], dofs[c["dof"]]]
        v = 0.0 if e < 0 else u[e]
        tol = c["rel_tol"] * abs(c["expected"]) if "rel_tol" in c else c["abs_tol"]
        assert abs(v - c["expected"]) <= tol, (c, v)


def test_field_formats_agree():
    m = decks.config2_hex_jitter(nx=6, ny=3, nz=3)
    m.xyz = np.round(m.xyz, 6)              # exact in 16-column large fields
    m.E = np.round(m.E, 3)
    a = deck.read_deck(decks.bdf_text(m, "free"), is_text=True)
    b = deck.read_deck(decks.bdf_text(m, "mixed"), is_text=True)
    same_model(a, b)
    assert np.array_equal(a.xyz, m.xyz) and np.array_equal(a.E, m.E)
    assert np.array_equal(merged_spc(a)[merged_spc(a) > 0], np.full(16, 7))


def test_nastran_numbers_and_cards():
    text = "\n".join([
        "SOL 101", "CEND",
        "SUBCASE 1", "  LOAD = 2", "  SPC = 5", "  STRESS(PLOT) = ALL",
        "SUBCASE 7", "  LOAD = 3", "  SPC = 5", "  DISPLACEMENT(SORT1) = ALL",
        "BEGIN BULK",
        "$ a comment line",
        "GRID,1,,1.5-3,-2.D-1,.5",
        "GRID,2,,7,1.E4,-1-2",
        "GRID    3               +3.+2   2.5     -.25",   # small field
        "GRID,4,,0.,1.,0.   $ trailing comment",
        "GRID,5,,1.,1.,1.",
        "GRID,6,,2.,0.,0.",
        "CTETRA,10,1,1,2,3,4",
        "CTETRA  11      1       2       3       4",        # blank-label continuation
        "        5",
        "PSOLID,1,9",
        "MAT1,9,2.1+5,,.3,7.85-9",
        "SPC1,5,123,1,THRU,3",
        "SPC1,5,45,6",
        "SPC,5,5,6,0.,6,1",
        "FORCE,2,4,0,2.0,1.,0.,-1.",
        "MOMENT,2,4,0,1.0,0.,3.,0.",
        "FORCE,3,5,0,1.0,0.,0.,1.",
        "GRAV,3,0,9.81,0.,0.,-1.",
        "PLOAD4,2,10,5.0,,,,THRU,11",
        "PLOAD4,3,2.5,11",
        "PARAM,POST,-1",
        "EIGRL,1,,,4",
        "ENDDATA",
    ]) + "\r\n"
    m = deck.read_deck(text, is_text=True)
    assert list(m.node_ids) == [1, 2, 3, 4, 5, 6]
    assert np.array_equal(m.xyz[:3], [[1.5e-3, -0.2, 0.5], [7.0, 1e4, -0.01], [300.0, 2.5, -0.25]])
    assert list(m.elem_ids) == [10, 11] and list(m.conn) == [0, 1, 2, 3, 1, 2, 3, 4]
    assert m.E[0] == 2.1e5 and m.nu[0] == 0.3 and m.rho[0] == 7.85e-9
    mask = merged_spc(m)
    assert list(mask) == [7, 7, 7, 0, 32, 25]   # SPC: node 5 C=6, node 6 C=1; SPC1 45 on 6
    assert m.subcase_ids == [1, 7]
    l1, l2 = m.loads
    assert np.array_equal(l1.nodal[3], [2.0, 0.0, -2.0, 0.0, 3.0, 0.0])
    assert list(l1.pload4.elem) == [0, 1] and np.all(l1.pload4.p == 5.0)
    assert l1.grav is None and np.allclose(l2.grav, [0.0, 0.0, -9.81])
    assert list(l2.pload4.elem) == [1] and np.all(l2.pload4.p == 2.5)
    assert l2.nodal[4, 2] == 1.0
    assert m.meta["unsupported_cards"] == ["EIGRL"]
    # PARAM POST is read (and ignored); SHELLFORM defaults to MITC4+
    assert m.shell_form == 0


def test_case_control_requests():
    from vibestran_amd import _lib as L
    import ctypes as C
    text = "\n".join(["SUBCASE 1", "  STRESS(PLOT) = ALL", "  DISPLACEMENT(SORT1) = ALL",
                      "SUBCASE 2", "  STRESS(SORT1) = ALL", "  DISPLACEMENT = ALL",
                      "SUBCASE 3", "  STRESS = NONE", "  DISPLACEMENT(PRINT,PLOT) = ALL",
                      "BEGIN BULK", "GRID,1,,0.,0.,0.", "ENDDATA"])
    lib = L.load()
    h = C.c_void_p()
    raw = text.encode()
    assert lib.vbh_deck_parse(raw, len(raw), C.byref(h)) == 0
    got = []
    for k in range(3):
        s = L.vbh_deck_subcase_info()
        assert lib.vbh_deck_subcase(h, k, C.byref(s)) == 0
        got.append((s.id, s.disp_print, s.stress_print))
    lib.vbh_deck_free(h)
    assert got == [(1, 0, 0), (2, 1, 1), (3, 1, 0)]


@pytest.mark.parametrize("card,msg", [
    ("GRID,1,2,0.,0.,0.", "Line 3: Node 1: CP=2 references undefined coordinate system"),
    ("GRID,1,,0.,0.,0.,4", "Line 3: Node 1: CD=4 references undefined coordinate system"),
    ("CORD2R,5,6,0.,0.,0.,0.,0.,1.,1.,0.,0.", "Line 3: Coordinate system 5: circular or unresolvable"),
    ("CORD2R,5,,0.,0.,0.,0.,0.,1.,0.,0.,2.", "Line 3: Coordinate system defining points are collinear"),
    ("MPC,1,1,7,1.0", "Line 3: MPC component 7 is not 1..6"),
    ("CHEXA,1,1,1,2,3,4,5,6,7", "Line 3: CHEXA has 7 nodes; expected 8"),
    ("GRID,1,,0.,abc,0.", "Line 3: cannot parse double: 'abc'"),
    ("CTETRA,1,8,1,1,1,1", "property 8 is not defined"),
    ("PLOAD4,1,1,1.0,,,,5", "PLOAD4 solid face selection requires both face grid fields"),
])
def test_errors_name_the_line(card, msg):
    text = "BEGIN BULK\nGRID,1,,0.,0.,0.\n" + card + "\nENDDATA\n"
    with pytest.raises(SolverError, match=msg):
        deck.read_deck(text, is_text=True)


def test_parallel_chunks_match_python_reader():
    # ~40k lines: cut into several chunks parsed in parallel
    m = decks.config2_hex_jitter(nx=40, ny=16, nz=16)
    text = decks.bdf_text(m, "free")
    assert text.count("\n") > 40000
    same_model(bdf.read_bdf(text, is_text=True), deck.read_deck(text, is_text=True))


def frame(o, ez, ex_hint):
    """Axes of build_axes for A = o, B = o + ez, C = o + ex_hint."""
    z = np.asarray(ez, float) / np.linalg.norm(ez)
    y = np.cross(z, np.asarray(ex_hint, float) / np.linalg.norm(ex_hint))
    y /= np.linalg.norm(y)
    x = np.cross(y, z)
    return np.asarray(o, float), np.stack([x / np.linalg.norm(x), y, z])


def test_coordinate_systems_resolve_to_basic():
    # CORD2C 10 in basic; CORD2R 20 defined in CORD2C 10 (RID chain);
    # CORD2S 30 in basic; CORD1R 40 from nodes 1, 2, 3.
    text = "\n".join([
        "BEGIN BULK",
        "CORD2C,10,,1.,2.,3.,1.,2.,4.,2.,2.,3.",
        "CORD2R,20,10,2.,90.,0.,2.,90.,1.,3.,90.,0.",
        "CORD2S,30,,0.,0.,0.,0.,0.,1.,1.,0.,0.",
        "CORD1R,40,1,2,3",
        "GRID,1,,0.,0.,0.",
        "GRID,2,,0.,0.,1.",
        "GRID,3,,0.,1.,0.",
        "GRID,4,10,2.,30.,0.5",
        "GRID,5,20,1.,0.,0.",
        "GRID,6,30,2.,90.,90.",
        "GRID,7,40,1.,0.,0.",
        "GRID,8,,5.,5.,5.,10",
        "CTETRA,1,1,1,2,3,4",
        "PSOLID,1,1", "MAT1,1,1.,,.3",
        "ENDDATA"]) + "\n"
    m = deck.read_deck(text, is_text=True)
    o10, ax10 = frame([1, 2, 3], [0, 0, 1], [1, 0, 0])
    th = np.radians(30.0)
    x4 = o10 + ax10.T @ np.array([2 * np.cos(th), 2 * np.sin(th), 0.5])
    # frame 20: A, B, C given as cylindrical points of frame 10
    pa = o10 + ax10.T @ np.array([0.0, 2.0, 0.0])
    pb = o10 + ax10.T @ np.array([0.0, 2.0, 1.0])
    pc = o10 + ax10.T @ np.array([0.0, 3.0, 0.0])
    o20, ax20 = frame(pa, pb - pa, pc - pa)
    x5 = o20 + ax20[0]
    x6 = np.array([0.0, 2.0, 0.0])            # rho 2, theta 90, phi 90
    x7 = np.array([0.0, 1.0, 0.0])            # CORD1R: z along node 2, x toward node 3
    idx = {int(g): i for i, g in enumerate(m.node_ids)}
    for nid, want in ((4, x4), (5, x5), (6, x6), (7, x7)):
        assert np.allclose(m.xyz[idx[nid]], want, rtol=0, atol=1e-13), (nid, m.xyz[idx[nid]], want)
    # node 8 keeps basic position, CD = 10: rotation columns e_r, e_theta, e_z
    assert list(m.cd) == [0, 0, 0, 0, 0, 0, 0, 10]
    d = np.array([5.0, 5.0, 5.0]) - o10
    t = np.arctan2(d[1], d[0])
    want = np.stack([[np.cos(t), np.sin(t), 0.0], [-np.sin(t), np.cos(t), 0.0], [0, 0, 1.0]], axis=1)
    assert np.allclose(m.cd_rot[idx[8]], want, atol=1e-14)
    assert np.allclose(m.cd_rot[idx[1]], np.eye(3))


def test_force_in_cid_frame_and_mpc_sets():
    text = "\n".join([
        "SUBCASE 1", "  LOAD = 1", "  SPC = 1", "  MPC = 3",
        "BEGIN BULK",
        "CORD2C,10,,0.,0.,0.,0.,0.,1.,1.,0.,0.",
        "GRID,1,,0.,0.,0.", "GRID,2,,1.,1.,0.", "GRID,3,,0.,1.,0.", "GRID,4,,0.,0.,1.",
        "CTETRA,1,1,1,2,3,4", "PSOLID,1,1", "MAT1,1,1.,,.3",
        "FORCE,1,2,10,2.,1.,0.,0.",                 # radial at 45 degrees
        "MOMENT,1,3,0,1.,0.,0.,1.",
        "MPC,1,2,1,1.,3,2,-.5",
        "MPC,2,4,3,2.,2,3,1.",
        "MPCADD,3,1,2",
        "MPC,3,1,1,1.,4,1,1.",                      # after the MPCADD: appended last
        "MPC,4,1,2,1.,4,2,1.",                      # other set
        "ENDDATA"]) + "\n"
    m = deck.read_deck(text, is_text=True)
    l = m.loads[0].nodal
    assert np.allclose(l[1, :3], [np.sqrt(2.0), np.sqrt(2.0), 0.0], atol=1e-14)
    assert np.array_equal(l[2], [0, 0, 0, 0, 0, 1.0])
    assert m.mpc.equations() == [[(1, 1, 1.0), (2, 2, -0.5)], [(3, 3, 2.0), (1, 3, 1.0)],
                                 [(0, 1, 1.0), (3, 1, 1.0)]]
    assert m.cd is None


def test_mpc_continuation_terms():
    text = "\n".join([
        "SUBCASE 1", "  MPC = 5",
        "BEGIN BULK",
        "GRID,1,,0.,0.,0.", "GRID,2,,1.,0.,0.", "GRID,3,,0.,1.,0.", "GRID,4,,0.,0.,1.",
        "CTETRA,1,1,1,2,3,4", "PSOLID,1,1", "MAT1,1,1.,,.3",
        "MPC     5       1       1       1.0     2       1       -.5",
        "+       3       1       .25     4       2       -.125",
        "ENDDATA"]) + "\n"
    m = deck.read_deck(text, is_text=True)
    # continuation fields 1..8 are appended after field 8 of the first line;
    # that blank field 8 is read as the next term's grid and ends the list
    # (process_mpc breaks on an empty G), so the reference keeps two terms
    assert m.mpc.equations() == [[(0, 1, 1.0), (1, 1, -0.5)]]
