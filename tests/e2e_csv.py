"""CSV side of the reference's e2e runner: read <stem>.node.csv/.elem.csv the
way tools/e2e_test_runner.py:193-240 does when pyNastran is absent
(csv.reader with skipinitialspace, '#' rows skipped, columns by position) and
apply an expected.json's node_displacement / element_stress checks with their
own tolerances (e2e_test_runner.py check semantics: rel_tol on |expected|, or
abs_tol)."""
import csv
import json
import os

NODE_DOFS = ("T1", "T2", "T3", "R1", "R2", "R3")
ELEM_COLS = ("sx", "sy", "sxy", "sz", "syz", "szx", "mx", "my", "mxy", "von_mises")
GOLD = os.path.join(os.path.dirname(__file__), "golden", "e2e")


def load_csv_results(node_csv, elem_csv):
    node_data, elem_data = {}, {}
    with open(node_csv, encoding="utf-8") as fh:
        for row in csv.reader(fh, skipinitialspace=True):
            if not row or row[0].startswith("#"):
                continue
            node_data[(int(row[0]), int(row[1]))] = {
                d: float(row[2 + i]) for i, d in enumerate(NODE_DOFS)}
    with open(elem_csv, encoding="utf-8") as fh:
        for row in csv.reader(fh, skipinitialspace=True):
            if not row or row[0].startswith("#"):
                continue
            elem_data[(int(row[0]), int(row[1]))] = {
                c: float(row[2 + i]) for i, c in enumerate(ELEM_COLS)}
    return node_data, elem_data


def uses_frames_or_mpc(name):
    """Whether a deck has coordinate systems or MPC cards (read only by the
    native reader)."""
    with open(os.path.join(GOLD, name + ".bdf")) as fh:
        return any(ln.upper().startswith(("MPC", "CORD")) for ln in fh)


def deck_checks(name):
    with open(os.path.join(GOLD, name + ".expected.json")) as fh:
        return json.load(fh)["checks"]


def _ok(val, c):
    if "rel_tol" in c:
        return abs(val - c["expected"]) <= c["rel_tol"] * abs(c["expected"])
    return abs(val - c["expected"]) <= c["abs_tol"]


def apply_checks(name, node_data, elem_data):
    """Every node_displacement and element_stress check of the deck must
    pass; returns how many ran."""
    ran = 0
    for c in deck_checks(name):
        if c["type"] == "node_displacement":
            val = node_data[(c["node_id"], c["subcase"])][c["dof"]]
        elif c["type"] == "element_stress":
            val = elem_data[(c["elem_id"], c["subcase"])][c["component"]]
        else:
            continue
        assert _ok(val, c), (name, c, val)
        ran += 1
    return ran
