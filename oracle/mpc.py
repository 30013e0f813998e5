"""CPU restatement of the reference's multi-point constraint elimination.

TEST INFRASTRUCTURE ONLY (same rule as the rest of ``oracle``): tests use it
as the checker of the HIP path's K_red = T^T K T, F_red = T^T F and
u_full = T u_red.

Follows src/core/mpc_handler.cpp function by function, in the reference's own
element-level form (the HIP path transforms the assembled operator instead):
  MpcOracle.__init__         MpcHandler::build (14-164)
  MpcOracle.t_column         MpcHandler::t_column (172-186)
  MpcOracle.apply_to_stiffness  MpcHandler::apply_to_stiffness (188-278): fast
                             remap when no element DOF is dependent, else the
                             compact dense T^T Ke T over the element's active
                             reduced columns
  MpcOracle.apply_to_force   MpcHandler::apply_to_force (280-310)
  MpcOracle.recover          MpcHandler::recover_dependent_dofs (312-330)
and the assembly loop of LinearStaticSolver::assemble
(src/solver/assembly_parallel.hpp:68-110) with the MPC transform.

Pinned by tests/test_oracle_mpc.py: against the Lagrange-multiplier solution
of the constrained problem (K u + C^T lambda = F, C u = 0, dense numpy solve)
on meshes with independent MPC equations, which fixes T's signs, indices, the
reduced numbering and the T^T K T / T^T F / T u algebra without restating
them; against hand-derived cases for the skip, cycle and chained-term rules;
and through the reference's cquad4_ring_* e2e fixtures.
"""
from __future__ import annotations

import numpy as np

from . import Csr, OracleError, SparseMatrixBuilder, hexa8_ke, penta6_ke, quad4_ke, tetra4_ke

CHEXA8, CPENTA6, CTETRA4, CQUAD4 = 1, 2, 3, 4


class MpcOracle:
    """MpcHandler over an equation numbering eq [nn, 6] (-1 constrained) of
    n_full equations; ``equations`` is a list of MPC equations, each a list
    of (node index, dof 1..6, coeff)."""

    def __init__(self, eq, n_full, equations, warn=None):
        self.n_full = int(n_full)
        self.eliminations = []          # [(dep_eq, [(eq, coeff), ...]), ...]
        self.skipped = 0
        eq = np.asarray(eq)
        for terms in equations:
            if not terms:
                continue
            dep_idx, max_abs = 0, 0.0
            for i, (_, _, c) in enumerate(terms):
                if abs(c) > max_abs:
                    max_abs, dep_idx = abs(c), i
            g, d, a_dep = terms[dep_idx]
            dep_eq = int(eq[g, d - 1])
            if dep_eq < 0:
                self.skipped += 1
                if warn is not None:
                    warn(terms)
                continue
            elim = []
            for i, (gi, di, ci) in enumerate(terms):
                if i == dep_idx:
                    continue
                q = int(eq[gi, di - 1])
                if q < 0:
                    continue
                elim.append((q, -ci / a_dep))
            self.eliminations.append((dep_eq, elim))
        self.index_map = list(range(self.n_full))
        self.n_red = self.n_full
        self.dep_to_elim = {}
        if not self.eliminations:
            return
        dep_set = {d for d, _ in self.eliminations}
        # cycle detection: Kahn over dependent -> dependent edges
        if any(q in dep_set for _, t in self.eliminations for q, _ in t):
            indeg = {d: 0 for d, _ in self.eliminations}
            fwd = {}
            for d, t in self.eliminations:
                for q, _ in t:
                    if q in dep_set:
                        fwd.setdefault(q, []).append(d)
                        indeg[d] += 1
            queue = [d for d, k in indeg.items() if k == 0]
            done = 0
            while queue:
                cur = queue.pop()
                done += 1
                for nx in fwd.get(cur, []):
                    indeg[nx] -= 1
                    if indeg[nx] == 0:
                        queue.append(nx)
            if done < len(indeg):
                raise OracleError("MPC dependency graph contains a cycle; "
                                  "check MPC equations for circular references")
        removed = 0
        for i in range(self.n_full):
            if i in dep_set:
                self.index_map[i] = -1
                removed += 1
            else:
                self.index_map[i] = i - removed
        self.n_red = self.n_full - removed
        # terms to reduced indices; terms on dependent DOFs dropped (146-156)
        self.eliminations = [(d, [(self.index_map[q], c) for q, c in t if self.index_map[q] >= 0])
                             for d, t in self.eliminations]
        for k, (d, _) in enumerate(self.eliminations):
            self.dep_to_elim[d] = k

    def t_column(self, full_eq):
        if full_eq < 0:
            return []
        k = self.dep_to_elim.get(full_eq)
        if k is not None:
            return self.eliminations[k][1]
        r = self.index_map[full_eq] if full_eq < self.n_full else -1
        return [] if r < 0 else [(r, 1.0)]

    def transform(self):
        """T as a dense [n_full, n_red] array."""
        T = np.zeros((self.n_full, self.n_red))
        for i in range(self.n_full):
            for r, c in self.t_column(i):
                T[i, r] += c
        return T

    def apply_to_stiffness(self, gdofs, ke, builder):
        gdofs = [int(g) for g in gdofs]
        ke = np.asarray(ke, dtype=np.float64).reshape(len(gdofs), len(gdofs))
        if not any(g >= 0 and g in self.dep_to_elim for g in gdofs):
            gd = [-1 if (g < 0 or g >= self.n_full) else self.index_map[g] for g in gdofs]
            builder.add_element_stiffness(gd, ke)
            return
        active = []
        rows = []
        for g in gdofs:
            ent = []
            for r, c in self.t_column(g):
                if r < 0:
                    continue
                if r not in active:
                    active.append(r)
                ent.append((active.index(r), c))
            rows.append(ent)
        na = len(active)
        if na == 0:
            return
        T = np.zeros((len(gdofs), na))
        for i, ent in enumerate(rows):
            for j, c in ent:
                T[i, j] += c
        builder.add_element_stiffness(active, T.T @ (ke @ T))

    def apply_to_force(self, gdofs, fe, F):
        for g, f in zip(gdofs, fe):
            for r, c in self.t_column(int(g)):
                if 0 <= r < len(F):
                    F[r] += c * f

    def recover(self, u_red):
        u = np.zeros(self.n_full)
        for i, r in enumerate(self.index_map):
            if r >= 0:
                u[i] = u_red[r]
        for d, t in self.eliminations:
            u[d] = sum(c * u_red[r] for r, c in t)
        return u


def element_ke(model, e):
    """Element stiffness and the element's node indices / DOFs per node."""
    nodes = model.conn[model.elem_ptr[e]:model.elem_ptr[e + 1]]
    x = model.xyz[nodes]
    t = int(model.elem_type[e])
    if t == CHEXA8:
        ke = hexa8_ke(x, model.E[e], model.nu[e], "SRI" if model.solid_form == 1 else "EAS")
    elif t == CPENTA6:
        ke = penta6_ke(x, model.E[e], model.nu[e])
    elif t == CTETRA4:
        ke = tetra4_ke(x, model.E[e], model.nu[e])
    elif t == CQUAD4:
        ke = quad4_ke(x, model.E[e], model.nu[e], model.thickness[e],
                      "MINDLIN" if model.shell_form == 1 else "MITC4")
    else:
        raise OracleError("unsupported element type")
    return ke, nodes, (6 if t == CQUAD4 else 3)


def assemble_model_mpc(model, eq, mpc: MpcOracle) -> Csr:
    """LinearStaticSolver::assemble with the MPC transform: every element's
    Ke through apply_to_stiffness into one builder of n_red equations."""
    b = SparseMatrixBuilder(mpc.n_red)
    for e in range(len(model.elem_type)):
        ke, nodes, dpn = element_ke(model, e)
        gd = np.asarray(eq)[nodes][:, :dpn].reshape(-1)
        mpc.apply_to_stiffness(gd, ke, b)
    return b.build_csr()


def nodal_rhs_mpc(model, eq, nodal, mpc: MpcOracle) -> np.ndarray:
    """apply_point_loads (linear_static.cpp:344-398): each node's six
    components through apply_to_force."""
    F = np.zeros(mpc.n_red)
    eq = np.asarray(eq)
    for i in range(model.nn):
        if np.any(nodal[i] != 0.0):
            mpc.apply_to_force(eq[i], nodal[i], F)
    return F
