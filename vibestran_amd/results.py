"""Result recovery on the GPU and the CSV result files.

Mirrors the reference's LinearStaticSolver::recover_results
(src/solver/linear_static.cpp:566-1201: node displacements, element centroid
and corner stresses) and CsvWriter::write (src/io/csv_writer.cpp:21-93), over
the mesh of the last ``assembly.assemble_model`` on a backend. The stresses
come from the HIP kernels of libvbh (vbh_recover_stress); the files from
libvbh's host writer (vbh_write_node_csv / vbh_write_elem_csv), which needs no
GPU.
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib as L
from .backend import _check
from .model import Model

# Element row columns, the element CSV's order after (elem_id, subcase_id).
STRESS_COLUMNS = ("sx", "sy", "sxy", "sz", "syz", "szx", "mx", "my", "mxy", "von_mises")
NODE_COLUMNS = ("T1", "T2", "T3", "R1", "R2", "R3")


def recover_stress(backend, model: Model, u, corners: bool = True):
    """Element stresses for solution ``u``: (centroid [ne, 10], corner
    [ne, 8, 10] or None), columns STRESS_COLUMNS. Corner rows follow the
    element's node order (CHEXA8 8, CPENTA6 6, CTETRA4 4, CQUAD4 4 rows; the
    rest are 0)."""
    lib = L.load()
    u = np.ascontiguousarray(u, dtype=np.float64)
    cen = np.empty((model.ne, 10), dtype=np.float64)
    cor = np.empty((model.ne, 8, 10), dtype=np.float64) if corners else None
    _check(lib.vbh_recover_stress(backend._h, L.dp(u), L.dp(cen),
                                  L.dp(cor) if corners else None))
    return cen, cor


def recover_displacements(backend, model: Model, u) -> np.ndarray:
    """Node displacements [nn, 6] (T1..R3) in model node order; constrained
    and inactive DOFs are 0."""
    lib = L.load()
    u = np.ascontiguousarray(u, dtype=np.float64)
    disp = np.empty((model.nn, 6), dtype=np.float64)
    _check(lib.vbh_recover_displacements(backend._h, L.dp(u), L.dp(disp)))
    return disp


def element_forces(backend, model: Model, u) -> np.ndarray:
    """Internal nodal forces sum_e K_e u_e, [nn, 6] in the basic frame, at every
    node DOF (constrained ones included)."""
    lib = L.load()
    u = np.ascontiguousarray(u, dtype=np.float64)
    f = np.empty((model.nn, 6), dtype=np.float64)
    _check(lib.vbh_element_forces(backend._h, L.dp(u), L.dp(f)))
    return f


def reactions(backend, model: Model, u, case) -> np.ndarray:
    """R = sum_e K_e u_e - F_applied, [nn, 6]: support reactions at
    constrained DOFs, minus the equilibrium residual at free DOFs (no
    reference counterpart)."""
    from .assembly import load_args, load_arrays
    lib = L.load()
    u = np.ascontiguousarray(u, dtype=np.float64)
    arrays = load_arrays(model, case)
    R = np.empty((model.nn, 6), dtype=np.float64)
    _check(lib.vbh_reactions(backend._h, L.dp(u), *load_args(arrays), L.dp(R)))
    return R


def csv_paths(stem: str):
    """<stem without extension>.node.csv / .elem.csv, as CsvWriter::write
    derives them (csv_writer.cpp:36-38, 60-62)."""
    base = os.path.splitext(str(stem))[0]
    return base + ".node.csv", base + ".elem.csv"


def write_csv(stem: str, model: Model, subcases):
    """Write the node and element CSV files for ``subcases``, an iterable of
    (subcase_id, disp [nn, 6] or None, centroid rows [ne, 10] or None); a None
    block is skipped like a subcase without DISP/STRESS output requests.
    Returns the two paths."""
    lib = L.load()
    node_path, elem_path = csv_paths(stem)
    ids = np.ascontiguousarray(model.node_ids, dtype=np.int32)
    eids = np.ascontiguousarray(model.element_ids(), dtype=np.int32)
    et = np.ascontiguousarray(model.elem_type, dtype=np.int32)
    subcases = list(subcases)
    # headers (and empty files when nothing is requested)
    _check(lib.vbh_write_node_csv(node_path.encode(), 0, 0, 0, None, None))
    _check(lib.vbh_write_elem_csv(elem_path.encode(), 0, 0, 0, None, None, None))
    for sid, disp, rows in subcases:
        if disp is not None:
            d = np.ascontiguousarray(disp, dtype=np.float64).reshape(model.nn, 6)
            _check(lib.vbh_write_node_csv(node_path.encode(), 1, int(sid), model.nn, L.ip(ids),
                                          L.dp(d)))
        if rows is not None:
            r = np.ascontiguousarray(rows, dtype=np.float64).reshape(model.ne, 10)
            _check(lib.vbh_write_elem_csv(elem_path.encode(), 1, int(sid), model.ne, L.ip(eids),
                                          L.ip(et), L.dp(r)))
    return node_path, elem_path
