"""Writes the golden fixtures of tests/golden: direct (scipy SuperLU) solutions
of the oracle-assembled systems, used to pin the oracle's PCG and the GPU path.
Run: python tests/golden/make_golden.py"""
import os
import sys

import numpy as np
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402
from vibestran_amd import decks  # noqa: E402
from vibestran_amd import model as M  # noqa: E402


def c1():
    m = decks.config1_hex_cantilever()
    eq, n = M.analysis_dof_map(m)
    K = O.assemble_solids(n, m.elem_type, m.elem_ptr, m.conn, m.xyz, eq, m.E, m.nu)
    F = M.assemble_rhs(m, eq, n, m.loads[0])
    u = spla.spsolve(K.to_scipy().tocsc(), F)
    rp, ci, v = K.arrays()
    np.savez_compressed(os.path.join(HERE, "c1_direct.npz"), n=n, u=u, F=F, nnz=len(ci),
                        row_ptr_sum=int(rp.sum()), col_sum=int(ci.astype(np.int64).sum()))


if __name__ == "__main__":
    c1()
    print("wrote golden fixtures")
