"""Reference e2e decks (tests/e2e/cases/*.bdf + *.expected.json, copied as
fixtures) solved through the CPU oracle: every node_displacement check of the
expected JSON must pass with the reference's own tolerances."""
import json
import os

import numpy as np
import pytest

import oracle as O
from e2e_csv import uses_frames_or_mpc
from vibestran_amd import bdf
from vibestran_amd import model as M

GOLD = os.path.join(os.path.dirname(__file__), "golden", "e2e")
# decks with coordinate systems or MPCs go through tests/test_oracle_mpc.py
DECKS = sorted(f[:-4] for f in os.listdir(GOLD) if f.endswith(".bdf") and not uses_frames_or_mpc(f[:-4]))
DOFS = {"T1": 0, "T2": 1, "T3": 2, "R1": 3, "R2": 4, "R3": 5}


def deck_checks(name):
    exp = json.load(open(os.path.join(GOLD, name + ".expected.json")))
    return [c for c in exp["checks"] if c["type"] == "node_displacement"]


def check_displacements(name, m, eq, u):
    idx = {int(g): i for i, g in enumerate(m.node_ids)}
    checks = deck_checks(name)
    assert checks
    for c in checks:
        e = eq[idx[c["node_id"]], DOFS[c["dof"]]]
        val = 0.0 if e < 0 else u[e]
        if "rel_tol" in c:
            assert abs(val - c["expected"]) <= c["rel_tol"] * abs(c["expected"]), (c, val)
        else:
            assert abs(val - c["expected"]) <= c["abs_tol"], (c, val)


@pytest.mark.parametrize("name", DECKS)
def test_e2e_deck_oracle(name):
    m = bdf.read_bdf(os.path.join(GOLD, name + ".bdf"))
    eq, n = M.analysis_dof_map(m)
    K = O.assemble_model(n, m, eq)
    F = M.assemble_rhs(m, eq, n, m.loads[0])
    u, it, rel = O.pcg_jacobi(K, F, tol=1e-12, max_iters=200000)
    check_displacements(name, m, eq, u)
