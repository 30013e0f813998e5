// oracle/src/oracle.cpp
//
// TEST INFRASTRUCTURE ONLY -- CPU FP64 restatement of the reference's
// linear-static hot path. Loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker; never by the product library.
//
// Pinning (what checks each function; see tests/test_oracle_*.py):
//   DofMap / analysis DOF map   test_dof_map.cpp:21-300 known answers + closed
//                               form (lexicographic rank of free pairs)
//   SparseMatrixBuilder/CsrData test_sparse_matrix.cpp:49-178 exact CSR and the
//                               symmetric matvec [9.5,-6.0]; brute-force dense
//   Element matrices            symmetry, rigid-body null space, PSD, patch
//                               tests (chexa8_axial_force deck, exact 1e-3),
//                               EAS == SRI... relations, Ke(E) linear in E,
//                               closed-form CTetra4 volume
//   Jacobi-PCG                  test_pcg_solver.cpp:80-222 closed forms and
//                               scipy direct solves on the same CSR
//   PLOAD4                      total force = p * face area, direction -n
#include "oracle.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// ---------------------------------------------------------------------------
// Small dense helpers (row-major)
// ---------------------------------------------------------------------------
struct M3 {
  double a[3][3];
};

// Eigen's 3x3 determinant: expansion along column 0
// (Eigen/src/LU/Determinant.h bruteforce_det3_helper).
double det3(const M3& m) {
  auto h = [&](int a, int b, int c) {
    return m.a[a][0] * (m.a[b][1] * m.a[c][2] - m.a[b][2] * m.a[c][1]);
  };
  return h(0, 1, 2) - h(1, 0, 2) + h(2, 0, 1);
}

M3 inv3(const M3& m) {
  M3 r;
  const double c00 = m.a[1][1] * m.a[2][2] - m.a[1][2] * m.a[2][1];
  const double c10 = m.a[1][2] * m.a[2][0] - m.a[1][0] * m.a[2][2];
  const double c20 = m.a[1][0] * m.a[2][1] - m.a[1][1] * m.a[2][0];
  const double det = m.a[0][0] * c00 + m.a[0][1] * c10 + m.a[0][2] * c20;
  const double id = 1.0 / det;
  r.a[0][0] = c00 * id;
  r.a[1][0] = c10 * id;
  r.a[2][0] = c20 * id;
  r.a[0][1] = (m.a[0][2] * m.a[2][1] - m.a[0][1] * m.a[2][2]) * id;
  r.a[1][1] = (m.a[0][0] * m.a[2][2] - m.a[0][2] * m.a[2][0]) * id;
  r.a[2][1] = (m.a[0][1] * m.a[2][0] - m.a[0][0] * m.a[2][1]) * id;
  r.a[0][2] = (m.a[0][1] * m.a[1][2] - m.a[0][2] * m.a[1][1]) * id;
  r.a[1][2] = (m.a[0][2] * m.a[1][0] - m.a[0][0] * m.a[1][2]) * id;
  r.a[2][2] = (m.a[0][0] * m.a[1][1] - m.a[0][1] * m.a[1][0]) * id;
  return r;
}

// isotropic_D: src/elements/solid_elements.cpp:17-27
void isotropic_D(double E, double nu, double D[6][6]) {
  const double lam = E * nu / ((1 + nu) * (1 - 2 * nu));
  const double mu = E / (2 * (1 + nu));
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) D[i][j] = 0.0;
  D[0][0] = lam + 2 * mu; D[0][1] = lam;          D[0][2] = lam;
  D[1][0] = lam;          D[1][1] = lam + 2 * mu; D[1][2] = lam;
  D[2][0] = lam;          D[2][1] = lam;          D[2][2] = lam + 2 * mu;
  D[3][3] = mu; D[4][4] = mu; D[5][5] = mu;
}

// acc(nr x nc) += A^T(nr x 6) * D * Bm(6 x nc) * w  -- the B^T D B update.
void add_BtDB(int nr, int nc, const double* A, const double D[6][6], const double* Bm,
              double w, double* acc) {
  std::vector<double> DB(6 * static_cast<size_t>(nc));
  for (int k = 0; k < 6; ++k)
    for (int j = 0; j < nc; ++j) {
      double s = 0.0;
      for (int l = 0; l < 6; ++l) s += D[k][l] * Bm[l * nc + j];
      DB[static_cast<size_t>(k * nc + j)] = s;
    }
  for (int i = 0; i < nr; ++i)
    for (int j = 0; j < nc; ++j) {
      double s = 0.0;
      for (int k = 0; k < 6; ++k) s += A[k * nr + i] * DB[static_cast<size_t>(k * nc + j)];
      acc[i * nc + j] += s * w;
    }
}

// Strain-displacement B (6 x 3*nn) from physical shape gradients, the layout
// of build_B in solid_elements.cpp:111-125 / 1085-1097.
void fill_B(int nn, const double* dnx, const double* dny, const double* dnz, double* B) {
  const int nc = 3 * nn;
  std::fill(B, B + 6 * nc, 0.0);
  for (int n = 0; n < nn; ++n) {
    const int c0 = 3 * n;
    B[0 * nc + c0 + 0] = dnx[n];
    B[1 * nc + c0 + 1] = dny[n];
    B[2 * nc + c0 + 2] = dnz[n];
    B[3 * nc + c0 + 0] = dny[n]; B[3 * nc + c0 + 1] = dnx[n];
    B[4 * nc + c0 + 1] = dnz[n]; B[4 * nc + c0 + 2] = dny[n];
    B[5 * nc + c0 + 0] = dnz[n]; B[5 * nc + c0 + 2] = dnx[n];
  }
}

// CHexa8::shape_functions: solid_elements.cpp:76-93
struct HexShape {
  double N[8], dxi[8], deta[8], dzeta[8];
};
HexShape hex_shape(double xi, double eta, double zeta) {
  static const double xn[8] = {-1, +1, +1, -1, -1, +1, +1, -1};
  static const double en[8] = {-1, -1, +1, +1, -1, -1, +1, +1};
  static const double zn[8] = {-1, -1, -1, -1, +1, +1, +1, +1};
  HexShape s;
  for (int i = 0; i < 8; ++i) {
    s.N[i] = 0.125 * (1 + xn[i] * xi) * (1 + en[i] * eta) * (1 + zn[i] * zeta);
    s.dxi[i] = 0.125 * xn[i] * (1 + en[i] * eta) * (1 + zn[i] * zeta);
    s.deta[i] = 0.125 * en[i] * (1 + xn[i] * xi) * (1 + zn[i] * zeta);
    s.dzeta[i] = 0.125 * zn[i] * (1 + xn[i] * xi) * (1 + en[i] * eta);
  }
  return s;
}

// Jacobian J(i,j) = sum_n dN_i/dparam * x_j (solid_elements.cpp:136-141)
M3 jacobian(int nn, const double* d0, const double* d1, const double* d2, const double* xyz) {
  M3 J{};
  for (int n = 0; n < nn; ++n) {
    const double x = xyz[3 * n], y = xyz[3 * n + 1], z = xyz[3 * n + 2];
    J.a[0][0] += d0[n] * x; J.a[0][1] += d0[n] * y; J.a[0][2] += d0[n] * z;
    J.a[1][0] += d1[n] * x; J.a[1][1] += d1[n] * y; J.a[1][2] += d1[n] * z;
    J.a[2][0] += d2[n] * x; J.a[2][1] += d2[n] * y; J.a[2][2] += d2[n] * z;
  }
  return J;
}

void phys_grad(int nn, const M3& Ji, const double* d0, const double* d1, const double* d2,
               double* dnx, double* dny, double* dnz) {
  for (int n = 0; n < nn; ++n) {
    dnx[n] = Ji.a[0][0] * d0[n] + Ji.a[0][1] * d1[n] + Ji.a[0][2] * d2[n];
    dny[n] = Ji.a[1][0] * d0[n] + Ji.a[1][1] * d1[n] + Ji.a[1][2] * d2[n];
    dnz[n] = Ji.a[2][0] * d0[n] + Ji.a[2][1] * d1[n] + Ji.a[2][2] * d2[n];
  }
}

// Fill column `col` of G (6 x 9) for the enhanced mode u = g * e_alpha
// (fill_G_col, solid_elements.cpp:1140-1151).
void fill_G_col(double* G, int col, double gx, double gy, double gz, int alpha) {
  if (alpha == 0) {
    G[0 * 9 + col] = gx; G[3 * 9 + col] = gy; G[5 * 9 + col] = gz;
  } else if (alpha == 1) {
    G[3 * 9 + col] = gx; G[1 * 9 + col] = gy; G[4 * 9 + col] = gz;
  } else {
    G[5 * 9 + col] = gx; G[4 * 9 + col] = gy; G[2 * 9 + col] = gz;
  }
}

// Ke = Kuu - Kua * Kaa^{-1} * Kua^T via 9x9 Cholesky (Eigen::LLT,
// solid_elements.cpp:1201-1204 / 1461-1465). Returns false when Kaa is not SPD.
bool condense(int nd, const double* Kuu, const double* Kua, const double* Kaa, double* Ke) {
  double L[9][9] = {};
  for (int j = 0; j < 9; ++j) {
    double d = Kaa[j * 9 + j];
    for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
    if (!(d > 0.0)) return false;
    L[j][j] = std::sqrt(d);
    for (int i = j + 1; i < 9; ++i) {
      double s = Kaa[i * 9 + j];
      for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
      L[i][j] = s / L[j][j];
    }
  }
  // X = Kaa^{-1} Kua^T  (9 x nd)
  std::vector<double> X(9 * static_cast<size_t>(nd));
  for (int c = 0; c < nd; ++c) {
    double y[9];
    for (int i = 0; i < 9; ++i) {
      double s = Kua[c * 9 + i];
      for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
      y[i] = s / L[i][i];
    }
    for (int i = 8; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < 9; ++k) s -= L[k][i] * X[static_cast<size_t>(k * nd + c)];
      X[static_cast<size_t>(i * nd + c)] = s / L[i][i];
    }
  }
  for (int i = 0; i < nd; ++i)
    for (int j = 0; j < nd; ++j) {
      double s = 0.0;
      for (int k = 0; k < 9; ++k) s += Kua[i * 9 + k] * X[static_cast<size_t>(k * nd + j)];
      Ke[i * nd + j] = Kuu[i * nd + j] - s;
    }
  return true;
}

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }
void orc_internal_set_error(const char* msg) { g_err = msg; }

// ===========================================================================
// DofMap (src/core/dof_map.cpp)
// ===========================================================================
struct OrcDofMap {
  std::unordered_map<int, std::array<int, 6>> blocks;
  int num_free = 0;
  int num_total = 0;
};

// DofMap::build: dof_map.cpp:10-33 (6 slots per node in sorted-ID order)
OrcDofMap* orc_dofmap_build(int nnodes, const int* node_ids, int dofs_per_node) {
  auto* m = new OrcDofMap();
  std::vector<int> ids(node_ids, node_ids + nnodes);
  std::sort(ids.begin(), ids.end());
  for (int nid : ids) {
    std::array<int, 6> eq;
    eq.fill(-1);
    for (int d = 0; d < dofs_per_node; ++d) eq[static_cast<size_t>(d)] = m->num_free++;
    m->blocks[nid] = eq;
    m->num_total += dofs_per_node;
  }
  m->num_total = m->num_free;
  return m;
}

void orc_dofmap_free(OrcDofMap* m) { delete m; }

// DofMap::constrain: dof_map.cpp:35-53
int orc_dofmap_constrain(OrcDofMap* m, int node, int dof0) {
  auto it = m->blocks.find(node);
  if (it == m->blocks.end()) return fail(ORC_ERR_ARG, "constrain: node not in DofMap");
  int& eq = it->second[static_cast<size_t>(dof0)];
  if (eq == -1) return ORC_OK;
  const int removed = eq;
  eq = -1;
  for (auto& kv : m->blocks)
    for (int& e : kv.second)
      if (e > removed) --e;
  --m->num_free;
  return ORC_OK;
}

// DofMap::constrain_batch: dof_map.cpp:55-97
int orc_dofmap_constrain_batch(OrcDofMap* m, int count, const int* nodes, const int* dofs) {
  std::vector<int> removed;
  removed.reserve(static_cast<size_t>(count));
  for (int i = 0; i < count; ++i) {
    auto it = m->blocks.find(nodes[i]);
    if (it == m->blocks.end()) return fail(ORC_ERR_ARG, "constrain_batch: node not in DofMap");
    const int eq = it->second[static_cast<size_t>(dofs[i])];
    if (eq != -1) removed.push_back(eq);
  }
  if (removed.empty()) return ORC_OK;
  std::sort(removed.begin(), removed.end());
  removed.erase(std::unique(removed.begin(), removed.end()), removed.end());
  for (auto& kv : m->blocks)
    for (int& e : kv.second) {
      if (e == -1) continue;
      e -= static_cast<int>(std::lower_bound(removed.begin(), removed.end(), e) - removed.begin());
    }
  for (int i = 0; i < count; ++i) m->blocks.at(nodes[i])[static_cast<size_t>(dofs[i])] = -1;
  m->num_free -= static_cast<int>(removed.size());
  return ORC_OK;
}

// DofMap::eq_index: dof_map.cpp:99-104 (unknown node -> CONSTRAINED_DOF)
int orc_dofmap_eq_index(const OrcDofMap* m, int node, int dof0) {
  auto it = m->blocks.find(node);
  if (it == m->blocks.end()) return -1;
  return it->second[static_cast<size_t>(dof0)];
}
int orc_dofmap_num_free(const OrcDofMap* m) { return m->num_free; }
int orc_dofmap_num_total(const OrcDofMap* m) { return m->num_total; }

// build_analysis_dof_map: analysis_support.cpp:176-214 with mark_element_dofs
// 39-68 (shells: all 6, solids: T1-T3). No RBE/MPC/CD frames in this subset.
int orc_analysis_dof_map(int nnodes, const int* node_ids, int nelem, const int* elem_type,
                         const int* elem_ptr, const int* elem_nodes, int nspc,
                         const int* spc_nodes, const int* spc_masks, int* eq_out) {
  OrcDofMap* m = orc_dofmap_build(nnodes, node_ids, 6);
  std::unordered_map<int, int> mask;  // node id -> active bits
  for (int i = 0; i < nnodes; ++i) mask[node_ids[i]] = 0;
  for (int e = 0; e < nelem; ++e) {
    const int bits = (elem_type[e] == ORC_CQUAD4) ? 0x3f : 0x07;
    for (int k = elem_ptr[e]; k < elem_ptr[e + 1]; ++k) mask[elem_nodes[k]] |= bits;
  }
  std::vector<int> cn, cd;
  for (const auto& kv : mask)
    for (int d = 0; d < 6; ++d)
      if (!(kv.second >> d & 1)) {
        cn.push_back(kv.first);
        cd.push_back(d);
      }
  int rc = orc_dofmap_constrain_batch(m, static_cast<int>(cn.size()), cn.data(), cd.data());
  if (rc != ORC_OK) { orc_dofmap_free(m); return -1; }
  cn.clear();
  cd.clear();
  for (int s = 0; s < nspc; ++s)
    for (int d = 0; d < 6; ++d)
      if (spc_masks[s] >> d & 1) {
        cn.push_back(spc_nodes[s]);
        cd.push_back(d);
      }
  rc = orc_dofmap_constrain_batch(m, static_cast<int>(cn.size()), cn.data(), cd.data());
  if (rc != ORC_OK) { orc_dofmap_free(m); return -1; }
  for (int i = 0; i < nnodes; ++i)
    for (int d = 0; d < 6; ++d) eq_out[6 * i + d] = orc_dofmap_eq_index(m, node_ids[i], d);
  const int nf = m->num_free;
  orc_dofmap_free(m);
  return nf;
}

// ===========================================================================
// Elements (src/elements/solid_elements.cpp)
// ===========================================================================

// CHexa8::stiffness_matrix (SRI): solid_elements.cpp:95-180
int orc_hexa8_ke_sri(const double* xyz, double E, double nu, double* Ke) {
  double D[6][6], Dvol[6][6] = {}, Ddev[6][6];
  isotropic_D(E, nu, D);
  const double lam = D[0][1];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Dvol[i][j] = lam;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) Ddev[i][j] = D[i][j] - Dvol[i][j];
  std::fill(Ke, Ke + 576, 0.0);
  const double gp = 1.0 / std::sqrt(3.0);
  const double g[2] = {-gp, gp};
  double B[6 * 24], dnx[8], dny[8], dnz[8];
  for (int gi = 0; gi < 2; ++gi)
    for (int gj = 0; gj < 2; ++gj)
      for (int gk = 0; gk < 2; ++gk) {
        HexShape s = hex_shape(g[gi], g[gj], g[gk]);
        M3 J = jacobian(8, s.dxi, s.deta, s.dzeta, xyz);
        const double detJ = det3(J);
        if (detJ <= 0) return fail(ORC_ERR_SOLVER, "CHEXA8: non-positive Jacobian");
        phys_grad(8, inv3(J), s.dxi, s.deta, s.dzeta, dnx, dny, dnz);
        fill_B(8, dnx, dny, dnz, B);
        add_BtDB(24, 24, B, Ddev, B, detJ, Ke);
      }
  HexShape s0 = hex_shape(0, 0, 0);
  M3 J0 = jacobian(8, s0.dxi, s0.deta, s0.dzeta, xyz);
  const double detJ0 = det3(J0);
  if (detJ0 <= 0) return fail(ORC_ERR_SOLVER, "CHEXA8: non-positive centroidal Jacobian");
  phys_grad(8, inv3(J0), s0.dxi, s0.deta, s0.dzeta, dnx, dny, dnz);
  fill_B(8, dnx, dny, dnz, B);
  add_BtDB(24, 24, B, Dvol, B, detJ0 * 8.0, Ke);
  return ORC_OK;
}

// CHexa8Eas::stiffness_matrix: solid_elements.cpp:1054-1206. The enhanced
// strain columns use the ROWS of J0^{-1} scaled by xi/eta/zeta (1131-1133).
int orc_hexa8_ke_eas(const double* xyz, double E, double nu, double* Ke) {
  double D[6][6];
  isotropic_D(E, nu, D);
  HexShape s0 = hex_shape(0, 0, 0);
  M3 J0 = jacobian(8, s0.dxi, s0.deta, s0.dzeta, xyz);
  const double detJ0 = det3(J0);
  if (detJ0 <= 0) return fail(ORC_ERR_SOLVER, "CHEXA8EAS: non-positive centroidal Jacobian");
  const M3 J0i = inv3(J0);
  std::vector<double> Kuu(576, 0.0), Kua(24 * 9, 0.0), Kaa(81, 0.0);
  const double gp = 1.0 / std::sqrt(3.0);
  const double g[2] = {-gp, gp};
  double B[6 * 24], G[6 * 9], dnx[8], dny[8], dnz[8];
  for (int gi = 0; gi < 2; ++gi)
    for (int gj = 0; gj < 2; ++gj)
      for (int gk = 0; gk < 2; ++gk) {
        const double xi = g[gi], eta = g[gj], zeta = g[gk];
        HexShape s = hex_shape(xi, eta, zeta);
        M3 J = jacobian(8, s.dxi, s.deta, s.dzeta, xyz);
        const double detJ = det3(J);
        if (detJ <= 0) return fail(ORC_ERR_SOLVER, "CHEXA8EAS: non-positive Jacobian");
        phys_grad(8, inv3(J), s.dxi, s.deta, s.dzeta, dnx, dny, dnz);
        fill_B(8, dnx, dny, dnz, B);
        std::fill(G, G + 54, 0.0);
        const double prm[3] = {xi, eta, zeta};
        for (int k = 0; k < 3; ++k) {
          const double gx = J0i.a[k][0] * prm[k], gy = J0i.a[k][1] * prm[k],
                       gz = J0i.a[k][2] * prm[k];
          for (int alpha = 0; alpha < 3; ++alpha) fill_G_col(G, 3 * k + alpha, gx, gy, gz, alpha);
        }
        const double scale = detJ0 / detJ;
        for (double& v : G) v *= scale;
        add_BtDB(24, 24, B, D, B, detJ, Kuu.data());
        add_BtDB(24, 9, B, D, G, detJ, Kua.data());
        add_BtDB(9, 9, G, D, G, detJ, Kaa.data());
      }
  if (!condense(24, Kuu.data(), Kua.data(), Kaa.data(), Ke))
    return fail(ORC_ERR_SOLVER, "CHEXA8EAS: Kaa factorization failed");
  return ORC_OK;
}

// CTetra4::stiffness_matrix: solid_elements.cpp:312-384. dN_i/dx = cof(1,i)/6V
// (cofactors of A4 = [1 x y z] rows), Ke = V B^T D B.
int orc_tetra4_ke(const double* xyz, double E, double nu, double* Ke) {
  double A4[4][4];
  for (int i = 0; i < 4; ++i) {
    A4[i][0] = 1.0;
    for (int d = 0; d < 3; ++d) A4[i][d + 1] = xyz[3 * i + d];
  }
  auto minor_det = [&](int skip_r, int skip_c) {
    M3 m;
    int ri = 0;
    for (int r = 0; r < 4; ++r) {
      if (r == skip_r) continue;
      int ci = 0;
      for (int c = 0; c < 4; ++c) {
        if (c == skip_c) continue;
        m.a[ri][ci++] = A4[r][c];
      }
      ++ri;
    }
    return det3(m);
  };
  double V6 = 0.0;
  for (int c = 0; c < 4; ++c) V6 += ((c % 2) ? -1.0 : 1.0) * A4[0][c] * minor_det(0, c);
  const double V = V6 / 6.0;
  if (V <= 0) return fail(ORC_ERR_SOLVER, "CTETRA4: non-positive volume");
  // cofA(i,j) = (-1)^(i+j) * det(minor removing row j, col i)
  double cof[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) cof[i][j] = (((i + j) % 2) ? -1.0 : 1.0) * minor_det(j, i);
  double dnx[4], dny[4], dnz[4], B[6 * 12], D[6][6];
  for (int n = 0; n < 4; ++n) {
    dnx[n] = cof[1][n] / V6;
    dny[n] = cof[2][n] / V6;
    dnz[n] = cof[3][n] / V6;
  }
  fill_B(4, dnx, dny, dnz, B);
  isotropic_D(E, nu, D);
  std::fill(Ke, Ke + 144, 0.0);
  add_BtDB(12, 12, B, D, B, V, Ke);
  return ORC_OK;
}

// CPenta6::shape_functions: solid_elements.cpp:510-548
struct PentaShape {
  double N[6], dL1[6], dL2[6], dz[6];
};
static PentaShape penta_shape(double L1, double L2, double zeta) {
  const double L3 = 1.0 - L1 - L2, zm = (1.0 - zeta) * 0.5, zp = (1.0 + zeta) * 0.5;
  PentaShape s;
  s.N[0] = L1 * zm; s.N[1] = L2 * zm; s.N[2] = L3 * zm;
  s.N[3] = L1 * zp; s.N[4] = L2 * zp; s.N[5] = L3 * zp;
  const double a[6] = {zm, 0.0, -zm, zp, 0.0, -zp};
  const double b[6] = {0.0, zm, -zm, 0.0, zp, -zp};
  const double c[6] = {-L1 * 0.5, -L2 * 0.5, -L3 * 0.5, L1 * 0.5, L2 * 0.5, L3 * 0.5};
  for (int i = 0; i < 6; ++i) {
    s.dL1[i] = a[i];
    s.dL2[i] = b[i];
    s.dz[i] = c[i];
  }
  return s;
}

// CPenta6Eas::stiffness_matrix: solid_elements.cpp:1333-1467. Enhancement
// functions (L1-1/3, L2-1/3, zeta) with gradients from the COLUMNS of J0^{-1}.
int orc_penta6_ke_eas(const double* xyz, double E, double nu, double* Ke) {
  double D[6][6];
  isotropic_D(E, nu, D);
  PentaShape s0 = penta_shape(1.0 / 3.0, 1.0 / 3.0, 0.0);
  M3 J0 = jacobian(6, s0.dL1, s0.dL2, s0.dz, xyz);
  const double detJ0 = det3(J0);
  if (detJ0 <= 0) return fail(ORC_ERR_SOLVER, "CPENTA6EAS: non-positive centroidal Jacobian");
  const M3 J0i = inv3(J0);
  std::vector<double> Kuu(324, 0.0), Kua(18 * 9, 0.0), Kaa(81, 0.0);
  const double tri[3][2] = {{2.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}, {1.0 / 6.0, 1.0 / 6.0}};
  const double gp = 1.0 / std::sqrt(3.0);
  const double ax[2] = {-gp, gp};
  double B[6 * 18], G[6 * 9], dnx[6], dny[6], dnz[6];
  for (int ti = 0; ti < 3; ++ti)
    for (int ai = 0; ai < 2; ++ai) {
      const double L1 = tri[ti][0], L2 = tri[ti][1], zeta = ax[ai];
      PentaShape s = penta_shape(L1, L2, zeta);
      M3 J = jacobian(6, s.dL1, s.dL2, s.dz, xyz);
      const double detJ = det3(J);
      if (detJ <= 0) return fail(ORC_ERR_SOLVER, "CPENTA6EAS: non-positive Jacobian");
      phys_grad(6, inv3(J), s.dL1, s.dL2, s.dz, dnx, dny, dnz);
      fill_B(6, dnx, dny, dnz, B);
      std::fill(G, G + 54, 0.0);
      const double f[3] = {L1 - 1.0 / 3.0, L2 - 1.0 / 3.0, zeta};
      for (int k = 0; k < 3; ++k) {
        const double gx = J0i.a[0][k] * f[k], gy = J0i.a[1][k] * f[k], gz = J0i.a[2][k] * f[k];
        for (int alpha = 0; alpha < 3; ++alpha) fill_G_col(G, 3 * k + alpha, gx, gy, gz, alpha);
      }
      const double scale = detJ0 / detJ;
      for (double& v : G) v *= scale;
      const double w = detJ * (1.0 / 6.0);
      add_BtDB(18, 18, B, D, B, w, Kuu.data());
      add_BtDB(18, 9, B, D, G, w, Kua.data());
      add_BtDB(9, 9, G, D, G, w, Kaa.data());
    }
  if (!condense(18, Kuu.data(), Kua.data(), Kaa.data(), Ke))
    return fail(ORC_ERR_SOLVER, "CPENTA6EAS: Kaa factorization failed");
  return ORC_OK;
}

// CHexa8(Eas)::mass_matrix: solid_elements.cpp:182-217, 1208-1243
int orc_hexa8_me(const double* xyz, double rho, double* Me) {
  std::fill(Me, Me + 576, 0.0);
  if (rho == 0.0) return ORC_OK;
  const double gp = 1.0 / std::sqrt(3.0);
  const double g[2] = {-gp, gp};
  for (int gi = 0; gi < 2; ++gi)
    for (int gj = 0; gj < 2; ++gj)
      for (int gk = 0; gk < 2; ++gk) {
        HexShape s = hex_shape(g[gi], g[gj], g[gk]);
        const double detJ = det3(jacobian(8, s.dxi, s.deta, s.dzeta, xyz));
        if (detJ <= 0) return fail(ORC_ERR_SOLVER, "CHEXA8: non-positive Jacobian in mass_matrix");
        for (int a = 0; a < 8; ++a)
          for (int b = 0; b < 8; ++b) {
            const double v = rho * s.N[a] * s.N[b] * detJ;
            for (int d = 0; d < 3; ++d) Me[(3 * a + d) * 24 + 3 * b + d] += v;
          }
      }
  return ORC_OK;
}

// CTetra4::mass_matrix: solid_elements.cpp:386-411
int orc_tetra4_me(const double* xyz, double rho, double* Me) {
  std::fill(Me, Me + 144, 0.0);
  if (rho == 0.0) return ORC_OK;
  double A4[4][4];
  for (int i = 0; i < 4; ++i) {
    A4[i][0] = 1.0;
    for (int d = 0; d < 3; ++d) A4[i][d + 1] = xyz[3 * i + d];
  }
  double V6 = 0.0;
  for (int c = 0; c < 4; ++c) {
    M3 m;
    for (int r = 1; r < 4; ++r) {
      int ci = 0;
      for (int cc = 0; cc < 4; ++cc)
        if (cc != c) m.a[r - 1][ci++] = A4[r][cc];
    }
    V6 += ((c % 2) ? -1.0 : 1.0) * A4[0][c] * det3(m);
  }
  const double V = std::fabs(V6) / 6.0;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      const double v = (a == b) ? rho * V / 10.0 : rho * V / 20.0;
      for (int d = 0; d < 3; ++d) Me[(3 * a + d) * 12 + 3 * b + d] = v;
    }
  return ORC_OK;
}

// CPenta6Eas::mass_matrix: solid_elements.cpp:1469-1509
int orc_penta6_me(const double* xyz, double rho, double* Me) {
  std::fill(Me, Me + 324, 0.0);
  if (rho == 0.0) return ORC_OK;
  const double tri[3][2] = {{2.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}, {1.0 / 6.0, 1.0 / 6.0}};
  const double ax[2] = {-1.0 / std::sqrt(3.0), 1.0 / std::sqrt(3.0)};
  for (int ti = 0; ti < 3; ++ti)
    for (int ai = 0; ai < 2; ++ai) {
      PentaShape s = penta_shape(tri[ti][0], tri[ti][1], ax[ai]);
      const double detJ = det3(jacobian(6, s.dL1, s.dL2, s.dz, xyz));
      if (detJ <= 0) return fail(ORC_ERR_SOLVER, "CPENTA6: non-positive Jacobian in mass_matrix");
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) {
          const double v = rho * s.N[a] * s.N[b] * detJ * (1.0 / 6.0);
          for (int d = 0; d < 3; ++d) Me[(3 * a + d) * 18 + 3 * b + d] += v;
        }
    }
  return ORC_OK;
}

// ===========================================================================
// SparseMatrixBuilder / CsrData (src/core/sparse_matrix.cpp)
// ===========================================================================
struct Trip {
  int row, col;
  double v;
};
struct OrcBuilder {
  int n;
  std::vector<Trip> t;
};
struct OrcCsr {
  int n = 0, nnz = 0;
  bool lower = false;
  std::vector<int> row_ptr, col_ind;
  std::vector<double> values;
};

OrcBuilder* orc_builder_create(int n) { return new OrcBuilder{n, {}}; }
void orc_builder_free(OrcBuilder* b) { delete b; }
void orc_builder_add(OrcBuilder* b, int row, int col, double v) { b->t.push_back({row, col, v}); }
long long orc_builder_num_triplets(const OrcBuilder* b) { return static_cast<long long>(b->t.size()); }

// add_element_stiffness: sparse_matrix.cpp:16-35 (lower triangle by global
// index, exact 0.0 entries dropped, constrained DOFs skipped)
int orc_builder_add_element(OrcBuilder* b, int n, const int* gd, const double* ke) {
  for (int i = 0; i < n; ++i) {
    if (gd[i] == -1) continue;
    for (int j = 0; j < n; ++j) {
      if (gd[j] == -1) continue;
      if (gd[i] < gd[j]) continue;
      const double v = ke[static_cast<size_t>(i * n + j)];
      if (v != 0.0) b->t.push_back({gd[i], gd[j], v});
    }
  }
  return ORC_OK;
}

// build_csr: sparse_matrix.cpp:62-150. Histogram -> scatter -> per-row sort by
// column -> dedup by summation. The reference's per-row std::sort is unstable,
// so its summation order for equal columns is unspecified; this restatement
// uses a stable sort (element order) to make the order defined.
OrcCsr* orc_builder_build_csr(OrcBuilder* b) {
  const int n = b->n;
  bool lower_only = true;
  for (const auto& t : b->t)
    if (t.row < t.col) { lower_only = false; break; }
  std::vector<int> bucket(static_cast<size_t>(n + 1), 0);
  for (const auto& t : b->t) ++bucket[static_cast<size_t>(t.row + 1)];
  std::partial_sum(bucket.begin(), bucket.end(), bucket.begin());
  std::vector<int> cursor(bucket.begin(), bucket.begin() + n);
  std::vector<Trip> staged(b->t.size());
  for (const auto& t : b->t) staged[static_cast<size_t>(cursor[static_cast<size_t>(t.row)]++)] = t;
  b->t.clear();
  b->t.shrink_to_fit();
  for (int r = 0; r < n; ++r)
    std::stable_sort(staged.begin() + bucket[static_cast<size_t>(r)],
                     staged.begin() + bucket[static_cast<size_t>(r + 1)],
                     [](const Trip& x, const Trip& y) { return x.col < y.col; });
  auto* c = new OrcCsr();
  c->n = n;
  c->row_ptr.assign(static_cast<size_t>(n + 1), 0);
  c->col_ind.resize(staged.size());
  c->values.resize(staged.size());
  int out = 0;
  for (int r = 0; r < n; ++r) {
    const int beg = bucket[static_cast<size_t>(r)], end = bucket[static_cast<size_t>(r + 1)];
    for (int i = beg; i < end; ++i) {
      if (i > beg && staged[static_cast<size_t>(i)].col == staged[static_cast<size_t>(i - 1)].col) {
        c->values[static_cast<size_t>(out - 1)] += staged[static_cast<size_t>(i)].v;
      } else {
        c->col_ind[static_cast<size_t>(out)] = staged[static_cast<size_t>(i)].col;
        c->values[static_cast<size_t>(out)] = staged[static_cast<size_t>(i)].v;
        ++out;
      }
    }
    c->row_ptr[static_cast<size_t>(r + 1)] = out;
  }
  c->nnz = out;
  c->col_ind.resize(static_cast<size_t>(out));
  c->values.resize(static_cast<size_t>(out));
  c->lower = lower_only;
  return c;
}

void orc_csr_free(OrcCsr* c) { delete c; }
int orc_csr_n(const OrcCsr* c) { return c->n; }
int orc_csr_nnz(const OrcCsr* c) { return c->nnz; }
int orc_csr_is_lower(const OrcCsr* c) { return c->lower ? 1 : 0; }
void orc_csr_copy(const OrcCsr* c, int* rp, int* ci, double* v) {
  if (rp) std::copy(c->row_ptr.begin(), c->row_ptr.end(), rp);
  if (ci) std::copy(c->col_ind.begin(), c->col_ind.end(), ci);
  if (v) std::copy(c->values.begin(), c->values.end(), v);
}

OrcCsr* orc_csr_from_arrays(int n, int nnz, const int* rp, const int* ci, const double* v,
                            int lower_only) {
  auto* c = new OrcCsr();
  c->n = n;
  c->nnz = nnz;
  c->lower = lower_only != 0;
  c->row_ptr.assign(rp, rp + n + 1);
  c->col_ind.assign(ci, ci + nnz);
  c->values.assign(v, v + nnz);
  return c;
}

// CsrData::multiply: sparse_matrix.cpp:152-181
int orc_csr_multiply(const OrcCsr* c, const double* x, double* y) {
  const int n = c->n;
  std::fill(y, y + n, 0.0);
  for (int r = 0; r < n; ++r)
    for (int k = c->row_ptr[static_cast<size_t>(r)]; k < c->row_ptr[static_cast<size_t>(r + 1)]; ++k) {
      const int col = c->col_ind[static_cast<size_t>(k)];
      const double v = c->values[static_cast<size_t>(k)];
      y[r] += v * x[col];
      if (c->lower && col != r) y[col] += v * x[r];
    }
  return ORC_OK;
}

// CsrData::expanded_symmetric: sparse_matrix.cpp:183-235
OrcCsr* orc_csr_expanded_symmetric(const OrcCsr* c) {
  auto* f = new OrcCsr();
  if (!c->lower) { *f = *c; return f; }
  const int n = c->n;
  f->n = n;
  f->lower = false;
  int off = 0;
  for (int r = 0; r < n; ++r)
    for (int k = c->row_ptr[static_cast<size_t>(r)]; k < c->row_ptr[static_cast<size_t>(r + 1)]; ++k)
      if (c->col_ind[static_cast<size_t>(k)] != r) ++off;
  f->nnz = c->nnz + off;
  f->row_ptr.assign(static_cast<size_t>(n + 1), 0);
  for (int r = 0; r < n; ++r)
    for (int k = c->row_ptr[static_cast<size_t>(r)]; k < c->row_ptr[static_cast<size_t>(r + 1)]; ++k) {
      const int col = c->col_ind[static_cast<size_t>(k)];
      ++f->row_ptr[static_cast<size_t>(r + 1)];
      if (col != r) ++f->row_ptr[static_cast<size_t>(col + 1)];
    }
  std::partial_sum(f->row_ptr.begin(), f->row_ptr.end(), f->row_ptr.begin());
  f->col_ind.resize(static_cast<size_t>(f->nnz));
  f->values.resize(static_cast<size_t>(f->nnz));
  std::vector<int> cur(f->row_ptr.begin(), f->row_ptr.begin() + n);
  for (int r = 0; r < n; ++r)
    for (int k = c->row_ptr[static_cast<size_t>(r)]; k < c->row_ptr[static_cast<size_t>(r + 1)]; ++k) {
      const int col = c->col_ind[static_cast<size_t>(k)];
      const double v = c->values[static_cast<size_t>(k)];
      int o = cur[static_cast<size_t>(r)]++;
      f->col_ind[static_cast<size_t>(o)] = col;
      f->values[static_cast<size_t>(o)] = v;
      if (col != r) {
        o = cur[static_cast<size_t>(col)]++;
        f->col_ind[static_cast<size_t>(o)] = r;
        f->values[static_cast<size_t>(o)] = v;
      }
    }
  return f;
}

// CsrData::lower_triangle: sparse_matrix.cpp:237-274
OrcCsr* orc_csr_lower_triangle(const OrcCsr* c) {
  auto* l = new OrcCsr();
  if (c->lower) { *l = *c; return l; }
  const int n = c->n;
  l->n = n;
  l->lower = true;
  l->row_ptr.assign(static_cast<size_t>(n + 1), 0);
  for (int r = 0; r < n; ++r) {
    for (int k = c->row_ptr[static_cast<size_t>(r)]; k < c->row_ptr[static_cast<size_t>(r + 1)]; ++k)
      if (c->col_ind[static_cast<size_t>(k)] <= r) {
        l->col_ind.push_back(c->col_ind[static_cast<size_t>(k)]);
        l->values.push_back(c->values[static_cast<size_t>(k)]);
      }
    l->row_ptr[static_cast<size_t>(r + 1)] = static_cast<int>(l->col_ind.size());
  }
  l->nnz = static_cast<int>(l->col_ind.size());
  return l;
}

// Mesh assembly through the element factory's default formulations
// (element_factory.cpp:11-104) and assemble_element_matrix's serial branch
// (assembly_parallel.hpp:101-108), with solid_global_dof_indices
// (solid_elements.cpp:29-42) and the shell's six DOFs per node (cquad4.cpp:855-864).
OrcCsr* orc_assemble_model(int n, int nelem, const int* et, const int* ep, const int* conn,
                           const double* xyz, const int* eq, const double* E, const double* nu,
                           const double* thick, int solid_form, int shell_form) {
  OrcBuilder b{n, {}};
  std::vector<double> ke(576), xe(24);
  std::vector<int> gd(24);
  for (int e = 0; e < nelem; ++e) {
    const int nn = ep[e + 1] - ep[e];
    const int dpn = (et[e] == ORC_CQUAD4) ? 6 : 3;
    for (int a = 0; a < nn; ++a) {
      const int node = conn[ep[e] + a];
      for (int d = 0; d < 3; ++d) xe[static_cast<size_t>(3 * a + d)] = xyz[3 * node + d];
      for (int d = 0; d < dpn; ++d) gd[static_cast<size_t>(dpn * a + d)] = eq[6 * node + d];
    }
    int rc;
    if (et[e] == ORC_CHEXA8)
      rc = solid_form == 1 ? orc_hexa8_ke_sri(xe.data(), E[e], nu[e], ke.data())
                           : orc_hexa8_ke_eas(xe.data(), E[e], nu[e], ke.data());
    else if (et[e] == ORC_CTETRA4)
      rc = orc_tetra4_ke(xe.data(), E[e], nu[e], ke.data());
    else if (et[e] == ORC_CPENTA6)
      rc = orc_penta6_ke_eas(xe.data(), E[e], nu[e], ke.data());
    else if (et[e] == ORC_CQUAD4)
      rc = orc_quad4_ke(xe.data(), E[e], nu[e], 0.0, thick[e], 1.0, 0.833333, shell_form,
                        ke.data());
    else
      rc = fail(ORC_ERR_ARG, "orc_assemble_model: unsupported element type");
    if (rc != ORC_OK) return nullptr;
    orc_builder_add_element(&b, dpn * nn, gd.data(), ke.data());
  }
  return orc_builder_build_csr(&b);
}

int orc_grav_rhs(int n, int nelem, const int* et, const int* ep, const int* conn,
                 const double* xyz, const int* eq, const double* rho, const double* thick,
                 const double* g, double* F) {
  std::fill(F, F + n, 0.0);
  std::vector<double> me(576), xe(24), ae(24), fe(24);
  for (int e = 0; e < nelem; ++e) {
    const int nn = ep[e + 1] - ep[e];
    const int dpn = (et[e] == ORC_CQUAD4) ? 6 : 3;
    const int nd = dpn * nn;
    for (int a = 0; a < nn; ++a) {
      const int node = conn[ep[e] + a];
      for (int d = 0; d < 3; ++d) xe[static_cast<size_t>(3 * a + d)] = xyz[3 * node + d];
      for (int d = 0; d < dpn; ++d) ae[static_cast<size_t>(dpn * a + d)] = (d < 3) ? g[d] : 0.0;
    }
    int rc;
    if (et[e] == ORC_CHEXA8) rc = orc_hexa8_me(xe.data(), rho[e], me.data());
    else if (et[e] == ORC_CTETRA4) rc = orc_tetra4_me(xe.data(), rho[e], me.data());
    else if (et[e] == ORC_CPENTA6) rc = orc_penta6_me(xe.data(), rho[e], me.data());
    else if (et[e] == ORC_CQUAD4) rc = orc_quad4_me(xe.data(), rho[e], thick[e], me.data());
    else rc = fail(ORC_ERR_ARG, "orc_grav_rhs: unsupported element type");
    if (rc != ORC_OK) return rc;
    double mmax = 0.0, fmax = 0.0;
    for (int i = 0; i < nd * nd; ++i) mmax = std::max(mmax, std::fabs(me[static_cast<size_t>(i)]));
    if (mmax < 1e-30) continue;  // linear_static.cpp:452-453
    for (int i = 0; i < nd; ++i) {
      double sv = 0.0;
      for (int j = 0; j < nd; ++j) sv += me[static_cast<size_t>(i * nd + j)] * ae[static_cast<size_t>(j)];
      fe[static_cast<size_t>(i)] = sv;
      fmax = std::max(fmax, std::fabs(sv));
    }
    if (fmax < 1e-30) continue;  // 495-497
    for (int a = 0; a < nn; ++a) {
      const int node = conn[ep[e] + a];
      for (int d = 0; d < dpn; ++d) {
        const int q = eq[6 * node + d];
        if (q >= 0) F[q] += fe[static_cast<size_t>(dpn * a + d)];
      }
    }
  }
  return ORC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Element stress recovery: LinearStaticSolver::recover_results, element part
// (linear_static.cpp:676-1197). Rows hold the CSV columns
// [sx, sy, sxy, sz, syz, szx, mx, my, mxy, von_mises] (csv_writer.cpp:66-85).
// ---------------------------------------------------------------------------
namespace {

// solid_von_mises: linear_static.cpp:140-151
double solid_vm(const double* s) {
  return std::sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[1] - s[2]) * (s[1] - s[2]) +
                          (s[2] - s[0]) * (s[2] - s[0]) +
                          6.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])));
}

// sigma = D (B ue) (solid_sigma_from_B 186-196, no thermal strain) as a row.
void solid_row(int nn, const double* dnx, const double* dny, const double* dnz, double E,
               double nu, const double* ue, double* row) {
  std::vector<double> B(6 * 3 * static_cast<size_t>(nn));
  fill_B(nn, dnx, dny, dnz, B.data());
  double D[6][6];
  isotropic_D(E, nu, D);
  double eps[6], sig[6];
  for (int i = 0; i < 6; ++i) {
    double s = 0.0;
    for (int j = 0; j < 3 * nn; ++j) s += B[static_cast<size_t>(i * 3 * nn + j)] * ue[j];
    eps[i] = s;
  }
  for (int i = 0; i < 6; ++i) {
    double s = 0.0;
    for (int j = 0; j < 6; ++j) s += D[i][j] * eps[j];
    sig[i] = s;
  }
  const double r[10] = {sig[0], sig[1], sig[3], sig[2], sig[4], sig[5], 0.0, 0.0, 0.0,
                        solid_vm(sig)};
  std::memcpy(row, r, sizeof r);
}

// eval_chexa8 / eval_cpenta6 (1058-1117, 996-1056): displacement B at a point.
void iso_point_row(int et, const double* xe, double E, double nu, const double* ue, double r0,
                   double r1, double r2, double* row) {
  double d0[8], d1[8], d2[8];
  int nn;
  if (et == ORC_CHEXA8) {
    const HexShape s = hex_shape(r0, r1, r2);
    nn = 8;
    std::memcpy(d0, s.dxi, sizeof s.dxi);
    std::memcpy(d1, s.deta, sizeof s.deta);
    std::memcpy(d2, s.dzeta, sizeof s.dzeta);
  } else {
    const PentaShape s = penta_shape(r0, r1, r2);
    nn = 6;
    std::memcpy(d0, s.dL1, sizeof s.dL1);
    std::memcpy(d1, s.dL2, sizeof s.dL2);
    std::memcpy(d2, s.dz, sizeof s.dz);
  }
  const M3 J = jacobian(nn, d0, d1, d2, xe);
  const M3 Ji = inv3(J);
  double dnx[8], dny[8], dnz[8];
  phys_grad(nn, Ji, d0, d1, d2, dnx, dny, dnz);
  solid_row(nn, dnx, dny, dnz, E, nu, ue, row);
}

// eval_ctetra4 (830-882): gradients of the CST shape functions from the
// cofactors of [1 x y z] divided by its determinant.
void tet_row(const double* xe, double E, double nu, const double* ue, double* row) {
  double A[4][4];
  for (int r = 0; r < 4; ++r) {
    A[r][0] = 1.0;
    for (int d = 0; d < 3; ++d) A[r][1 + d] = xe[3 * r + d];
  }
  auto minor_det = [&](int skip_r, int skip_c) {
    M3 m;
    int ri = 0;
    for (int r = 0; r < 4; ++r) {
      if (r == skip_r) continue;
      int ci = 0;
      for (int c = 0; c < 4; ++c) {
        if (c == skip_c) continue;
        m.a[ri][ci++] = A[r][c];
      }
      ++ri;
    }
    return det3(m);
  };
  double V6 = 0.0;
  for (int c = 0; c < 4; ++c) V6 += ((c & 1) ? -1.0 : 1.0) * A[0][c] * minor_det(0, c);
  double dnx[4], dny[4], dnz[4];
  for (int n = 0; n < 4; ++n) {
    // cofA(i, n) = (-1)^(i+n) det(A without row n and column i)
    const double sg = (n & 1) ? -1.0 : 1.0;
    dnx[n] = -sg * minor_det(n, 1) / V6;
    dny[n] = sg * minor_det(n, 2) / V6;
    dnz[n] = -sg * minor_det(n, 3) / V6;
  }
  solid_row(4, dnx, dny, dnz, E, nu, ue, row);
}

// make_plate_stress_point (168-184) from orc_quad4_recover's six values.
void plate_row(const double* v6, double* row) {
  const double sx = v6[0], sy = v6[1], sxy = v6[2];
  const double r[10] = {sx, sy, sxy, 0.0, 0.0, 0.0, v6[3], v6[4], v6[5],
                        std::sqrt(sx * sx - sx * sy + sy * sy + 3.0 * sxy * sxy)};
  std::memcpy(row, r, sizeof r);
}

}  // namespace

extern "C" {

int orc_recover_stress(int nelem, const int* et, const int* ep, const int* conn,
                       const double* xyz, const int* eq, const double* E, const double* nu,
                       const double* thick, int shell_form, const double* u, double* centroid,
                       double* corner) {
  static const double kHex[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                    {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  static const double kPenta[6][3] = {{1, 0, -1}, {0, 1, -1}, {0, 0, -1},
                                      {1, 0, 1},  {0, 1, 1},  {0, 0, 1}};
  static const double kQuad[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
  double xe[24], ue[24];
  for (int e = 0; e < nelem; ++e) {
    const int nn = ep[e + 1] - ep[e];
    const int dpn = (et[e] == ORC_CQUAD4) ? 6 : 3;
    for (int a = 0; a < nn; ++a) {
      const int node = conn[ep[e] + a];
      for (int d = 0; d < 3; ++d) xe[3 * a + d] = xyz[3 * node + d];
      for (int d = 0; d < dpn; ++d) {
        const int q = eq[6 * node + d];
        ue[dpn * a + d] = q >= 0 ? u[q] : 0.0;
      }
    }
    double* cen = centroid + 10 * static_cast<size_t>(e);
    double* cor = corner ? corner + 80 * static_cast<size_t>(e) : nullptr;
    if (cor) std::fill(cor, cor + 80, 0.0);
    if (et[e] == ORC_CHEXA8) {
      iso_point_row(et[e], xe, E[e], nu[e], ue, 0.0, 0.0, 0.0, cen);
      if (cor)
        for (int c = 0; c < 8; ++c)
          iso_point_row(et[e], xe, E[e], nu[e], ue, kHex[c][0], kHex[c][1], kHex[c][2],
                        cor + 10 * c);
    } else if (et[e] == ORC_CPENTA6) {
      iso_point_row(et[e], xe, E[e], nu[e], ue, 1.0 / 3.0, 1.0 / 3.0, 0.0, cen);
      if (cor)
        for (int c = 0; c < 6; ++c)
          iso_point_row(et[e], xe, E[e], nu[e], ue, kPenta[c][0], kPenta[c][1], kPenta[c][2],
                        cor + 10 * c);
    } else if (et[e] == ORC_CTETRA4) {
      tet_row(xe, E[e], nu[e], ue, cen);
      if (cor)
        for (int c = 0; c < 4; ++c) std::memcpy(cor + 10 * c, cen, 10 * sizeof(double));
    } else if (et[e] == ORC_CQUAD4) {
      double v6[6];
      int rc = orc_quad4_recover(xe, E[e], nu[e], thick[e], 1.0, shell_form, ue, 0.0, 0.0, v6);
      if (rc != ORC_OK) return rc;
      plate_row(v6, cen);
      if (cor)
        for (int c = 0; c < 4; ++c) {
          rc = orc_quad4_recover(xe, E[e], nu[e], thick[e], 1.0, shell_form, ue, kQuad[c][0],
                                 kQuad[c][1], v6);
          if (rc != ORC_OK) return rc;
          plate_row(v6, cor + 10 * c);
        }
    } else {
      return fail(ORC_ERR_ARG, "orc_recover_stress: unsupported element type");
    }
  }
  return ORC_OK;
}

int orc_element_sensitivity(int n, int nelem, const int* et, const int* ep, const int* conn,
                            const double* xyz, const int* eq, const double* E, const double* nu,
                            const double* thick, int solid_form, int shell_form, int param,
                            const double* u, const double* v, double* out) {
  (void)n;
  std::vector<double> ke(576), xe(24);
  std::vector<int> gd(24);
  for (int e = 0; e < nelem; ++e) {
    const int nn = ep[e + 1] - ep[e];
    const int dpn = (et[e] == ORC_CQUAD4) ? 6 : 3;
    for (int a = 0; a < nn; ++a) {
      const int node = conn[ep[e] + a];
      for (int d = 0; d < 3; ++d) xe[static_cast<size_t>(3 * a + d)] = xyz[3 * node + d];
      for (int d = 0; d < dpn; ++d) gd[static_cast<size_t>(dpn * a + d)] = eq[6 * node + d];
    }
    int rc = ORC_OK;
    double scale = 1.0;
    if (param == 1 && et[e] != ORC_CQUAD4) {
      out[e] = 0.0;
      continue;
    }
    if (et[e] == ORC_CHEXA8)
      rc = solid_form == 1 ? orc_hexa8_ke_sri(xe.data(), E[e], nu[e], ke.data())
                           : orc_hexa8_ke_eas(xe.data(), E[e], nu[e], ke.data());
    else if (et[e] == ORC_CTETRA4)
      rc = orc_tetra4_ke(xe.data(), E[e], nu[e], ke.data());
    else if (et[e] == ORC_CPENTA6)
      rc = orc_penta6_ke_eas(xe.data(), E[e], nu[e], ke.data());
    else if (et[e] == ORC_CQUAD4 && param == 0)
      rc = orc_quad4_ke(xe.data(), E[e], nu[e], 0.0, thick[e], 1.0, 0.833333, shell_form, ke.data());
    else if (et[e] == ORC_CQUAD4)
      rc = orc_quad4_dke_dt(xe.data(), E[e], nu[e], 0.0, thick[e], 1.0, 0.833333, shell_form,
                            ke.data());
    else
      rc = fail(ORC_ERR_ARG, "orc_element_sensitivity: unsupported element type");
    if (rc != ORC_OK) return rc;
    if (param == 0) scale = 1.0 / E[e];  // K_e is linear in E
    const int nd = dpn * nn;
    double sum = 0.0;
    for (int i = 0; i < nd; ++i) {
      const int gi = gd[static_cast<size_t>(i)];
      if (gi < 0) continue;
      for (int j = 0; j < nd; ++j) {
        const int gj = gd[static_cast<size_t>(j)];
        if (gj < 0) continue;
        sum += u[gi] * ke[static_cast<size_t>(i * nd + j)] * v[gj];
      }
    }
    out[e] = -sum * scale;
  }
  return ORC_OK;
}

int orc_element_forces(int nelem, const int* et, const int* ep, const int* conn,
                       const double* xyz, const int* eq, const double* E, const double* nu,
                       const double* thick, int solid_form, int shell_form, const double* u,
                       int nnodes, double* f) {
  std::fill(f, f + 6 * static_cast<size_t>(nnodes), 0.0);
  std::vector<double> ke(576), xe(24), ue(24);
  for (int e = 0; e < nelem; ++e) {
    const int nn = ep[e + 1] - ep[e];
    const int dpn = (et[e] == ORC_CQUAD4) ? 6 : 3;
    const int nd = dpn * nn;
    for (int a = 0; a < nn; ++a) {
      const int node = conn[ep[e] + a];
      for (int d = 0; d < 3; ++d) xe[static_cast<size_t>(3 * a + d)] = xyz[3 * node + d];
      for (int d = 0; d < dpn; ++d) {
        const int q = eq[6 * node + d];
        ue[static_cast<size_t>(dpn * a + d)] = q >= 0 ? u[q] : 0.0;
      }
    }
    int rc;
    if (et[e] == ORC_CHEXA8)
      rc = solid_form == 1 ? orc_hexa8_ke_sri(xe.data(), E[e], nu[e], ke.data())
                           : orc_hexa8_ke_eas(xe.data(), E[e], nu[e], ke.data());
    else if (et[e] == ORC_CTETRA4)
      rc = orc_tetra4_ke(xe.data(), E[e], nu[e], ke.data());
    else if (et[e] == ORC_CPENTA6)
      rc = orc_penta6_ke_eas(xe.data(), E[e], nu[e], ke.data());
    else if (et[e] == ORC_CQUAD4)
      rc = orc_quad4_ke(xe.data(), E[e], nu[e], 0.0, thick[e], 1.0, 0.833333, shell_form, ke.data());
    else
      rc = fail(ORC_ERR_ARG, "orc_element_forces: unsupported element type");
    if (rc != ORC_OK) return rc;
    for (int i = 0; i < nd; ++i) {
      double sv = 0.0;
      for (int j = 0; j < nd; ++j) sv += ke[static_cast<size_t>(i * nd + j)] * ue[static_cast<size_t>(j)];
      f[6 * static_cast<size_t>(conn[ep[e] + i / dpn]) + i % dpn] += sv;
    }
  }
  return ORC_OK;
}

OrcCsr* orc_assemble_solids(int n, int nelem, const int* et, const int* ep, const int* conn,
                            const double* xyz, const int* eq, const double* E, const double* nu,
                            int formulation) {
  return orc_assemble_model(n, nelem, et, ep, conn, xyz, eq, E, nu, nullptr, formulation, 0);
}

// ===========================================================================
// Jacobi-PCG
// ===========================================================================
namespace {
struct FullCsr {
  int n;
  std::vector<int> rp, ci;
  std::vector<double> v;
};

FullCsr make_full(const OrcCsr* K) {
  OrcCsr* f = orc_csr_expanded_symmetric(K);
  FullCsr F{f->n, f->row_ptr, f->col_ind, f->values};
  delete f;
  return F;
}

void spmv(const FullCsr& A, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int r = 0; r < A.n; ++r) {
    double s = 0.0;
    for (int k = A.rp[static_cast<size_t>(r)]; k < A.rp[static_cast<size_t>(r + 1)]; ++k)
      s += A.v[static_cast<size_t>(k)] * x[A.ci[static_cast<size_t>(k)]];
    y[r] = s;
  }
}

double dot(int n, const double* a, const double* b) {
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// build_jacobi_diagonal: cuda_pcg_backend_impl.hpp:2076-2101
int jacobi_diag(const OrcCsr* K, std::vector<double>& d) {
  d.assign(static_cast<size_t>(K->n), 0.0);
  for (int i = 0; i < K->n; ++i) {
    bool found = false;
    for (int k = K->row_ptr[static_cast<size_t>(i)]; k < K->row_ptr[static_cast<size_t>(i + 1)]; ++k)
      if (K->col_ind[static_cast<size_t>(k)] == i) {
        const double v = K->values[static_cast<size_t>(k)];
        if (v == 0.0)
          return fail(ORC_ERR_SOLVER, "PCG: zero diagonal at row " + std::to_string(i) +
                                          " -- matrix is singular. Check boundary conditions.");
        d[static_cast<size_t>(i)] = v;
        found = true;
        break;
      }
    if (!found)
      return fail(ORC_ERR_SOLVER, "PCG: missing diagonal at row " + std::to_string(i) +
                                      " -- matrix is singular. Check boundary conditions.");
  }
  return ORC_OK;
}

struct ThreadScope {
  int saved = 1;
  explicit ThreadScope(int nt) {
#ifdef _OPENMP
    saved = omp_get_max_threads();
    if (nt > 0) omp_set_num_threads(nt);
#else
    (void)nt;
#endif
  }
  ~ThreadScope() {
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
  }
};
}  // namespace

// Jacobi-PCG recurrence: run_pcg_prepared (cuda_pcg_backend_impl.hpp:1795-1999)
// with Jacobi apply 629-639, scalars 675-738, update 740-751.
int orc_pcg_jacobi(const OrcCsr* K, const double* Fv, double tol, int max_iters, int nthreads,
                   double* u, int* iters_out, double* rel_out) {
  ThreadScope ts(nthreads);
  const int n = K->n;
  if (n == 0) return fail(ORC_ERR_SOLVER, "PCG: stiffness matrix is empty -- no free DOFs");
  std::vector<double> d;
  if (jacobi_diag(K, d) != ORC_OK) return ORC_ERR_SOLVER;
  FullCsr A = make_full(K);
  std::vector<double> r(Fv, Fv + n), z(static_cast<size_t>(n)), p(static_cast<size_t>(n)),
      Ap(static_cast<size_t>(n));
  std::fill(u, u + n, 0.0);
  const double rr0 = dot(n, r.data(), r.data());
  const double norm_b = std::sqrt(std::fabs(rr0));
  if (norm_b < 1e-300) {
    *iters_out = 0;
    *rel_out = 0.0;
    return ORC_OK;
  }
  for (int i = 0; i < n; ++i) z[static_cast<size_t>(i)] = r[static_cast<size_t>(i)] / d[static_cast<size_t>(i)];
  p = z;
  double rz = dot(n, r.data(), z.data());
  if (rz <= 0.0)
    return fail(ORC_ERR_SOLVER, "PCG: non-positive initial r*z=" + std::to_string(rz) +
                                    " -- preconditioner is not SPD or matrix is singular");
  int iter = 0;
  double rel = 1.0;
  for (; iter < max_iters; ++iter) {
    spmv(A, p.data(), Ap.data());
    const double pAp = dot(n, p.data(), Ap.data());
    if (!(pAp > 0.0)) {  // breakdown: reported as non-convergence (1934-1942)
      rel = 1e300;
      iter = max_iters;
      break;
    }
    const double alpha = rz / pAp;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      u[i] += alpha * p[static_cast<size_t>(i)];
      r[static_cast<size_t>(i)] -= alpha * Ap[static_cast<size_t>(i)];
      z[static_cast<size_t>(i)] = r[static_cast<size_t>(i)] / d[static_cast<size_t>(i)];
    }
    const double rz_new = dot(n, r.data(), z.data());
    const double rr = dot(n, r.data(), r.data());
    rel = std::sqrt(std::fabs(rr)) / norm_b;
    if (rel < tol) {
      ++iter;
      break;
    }
    const double beta = rz_new / rz;
    rz = rz_new;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) p[static_cast<size_t>(i)] = z[static_cast<size_t>(i)] + beta * p[static_cast<size_t>(i)];
  }
  *iters_out = iter;
  *rel_out = rel;
  if (iter >= max_iters) {
    char buf[256];
    std::snprintf(buf, sizeof buf,
                  "PCG: did not converge after %d iterations (relative residual %g > tolerance %g).",
                  max_iters, rel, tol);
    return fail(ORC_ERR_SOLVER, buf);
  }
  return ORC_OK;
}

int orc_pcg_jacobi_fixed(const OrcCsr* K, const double* Fv, int iters, int nthreads, double* u,
                         double* seconds) {
  ThreadScope ts(nthreads);
  const int n = K->n;
  std::vector<double> d;
  if (jacobi_diag(K, d) != ORC_OK) return ORC_ERR_SOLVER;
  FullCsr A = make_full(K);
  std::vector<double> r(Fv, Fv + n), z(static_cast<size_t>(n)), p(static_cast<size_t>(n)),
      Ap(static_cast<size_t>(n));
  std::fill(u, u + n, 0.0);
  for (int i = 0; i < n; ++i) z[static_cast<size_t>(i)] = r[static_cast<size_t>(i)] / d[static_cast<size_t>(i)];
  p = z;
  double rz = dot(n, r.data(), z.data());
  const auto t0 = std::chrono::steady_clock::now();
  for (int it = 0; it < iters; ++it) {
    spmv(A, p.data(), Ap.data());
    const double pAp = dot(n, p.data(), Ap.data());
    const double alpha = rz / pAp;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      u[i] += alpha * p[static_cast<size_t>(i)];
      r[static_cast<size_t>(i)] -= alpha * Ap[static_cast<size_t>(i)];
      z[static_cast<size_t>(i)] = r[static_cast<size_t>(i)] / d[static_cast<size_t>(i)];
    }
    const double rz_new = dot(n, r.data(), z.data());
    volatile double rr = dot(n, r.data(), r.data());
    (void)rr;
    const double beta = rz_new / rz;
    rz = rz_new;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) p[static_cast<size_t>(i)] = z[static_cast<size_t>(i)] + beta * p[static_cast<size_t>(i)];
  }
  *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return ORC_OK;
}

// compute_true_relative_residual: cuda_pcg_backend_impl.hpp:90-114
double orc_true_relative_residual(const OrcCsr* K, const double* u, const double* Fv) {
  std::vector<double> Ku(static_cast<size_t>(K->n));
  orc_csr_multiply(K, u, Ku.data());
  double rn = 0.0, bn = 0.0;
  for (int i = 0; i < K->n; ++i) {
    const double ri = Fv[i] - Ku[static_cast<size_t>(i)];
    rn += ri * ri;
    bn += Fv[i] * Fv[i];
  }
  return bn > 1e-300 ? std::sqrt(rn / bn) : std::sqrt(rn);
}

// ===========================================================================
// PLOAD4 on a CHEXA8 face (linear_static_pressure.cpp)
// ===========================================================================
int orc_pload4_hexa8(const double* xyz, int g1, int g34, const double* pr, double* f24) {
  static const int faces[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                  {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};
  auto P = [&](int i, int d) { return xyz[3 * i + d]; };
  // orient_face_outward: 420-428
  auto orient = [&](const int* f, int* o) {
    double e1[3], e2[3], nrm[3], cen[3] = {0, 0, 0}, fc[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) {
      e1[d] = P(f[1], d) - P(f[0], d);
      e2[d] = P(f[2], d) - P(f[0], d);
    }
    nrm[0] = e1[1] * e2[2] - e1[2] * e2[1];
    nrm[1] = e1[2] * e2[0] - e1[0] * e2[2];
    nrm[2] = e1[0] * e2[1] - e1[1] * e2[0];
    for (int i = 0; i < 8; ++i)
      for (int d = 0; d < 3; ++d) cen[d] += P(i, d);
    for (int i = 0; i < 4; ++i)
      for (int d = 0; d < 3; ++d) fc[d] += P(f[i], d);
    double dotv = 0.0;
    for (int d = 0; d < 3; ++d) dotv += nrm[d] * (cen[d] / 8.0 - fc[d] / 4.0);
    if (dotv > 0.0) {
      o[0] = f[0]; o[1] = f[3]; o[2] = f[2]; o[3] = f[1];
    } else {
      for (int i = 0; i < 4; ++i) o[i] = f[i];
    }
  };
  int face[4];
  if (g1 < 0 && g34 < 0) {
    orient(faces[0], face);
  } else {
    bool found = false;
    for (int fi = 0; fi < 6 && !found; ++fi) {
      bool h1 = false, h34 = false;
      for (int i = 0; i < 4; ++i) {
        h1 = h1 || faces[fi][i] == g1;
        h34 = h34 || faces[fi][i] == g34;
      }
      if (!h1 || !h34) continue;
      int o[4];
      orient(faces[fi], o);
      int pos = 0;
      while (o[pos] != g1) ++pos;
      int rot[4];
      for (int i = 0; i < 4; ++i) rot[i] = o[(pos + i) % 4];
      if (rot[2] == g34) {
        for (int i = 0; i < 4; ++i) face[i] = rot[i];
        found = true;
      }
    }
    if (!found) return fail(ORC_ERR_SOLVER, "PLOAD4 could not resolve face from G1/G34");
  }
  // integrate_quad_surface_load: 154-196, normal_sign = -1
  std::fill(f24, f24 + 24, 0.0);
  const double gp = 1.0 / std::sqrt(3.0);
  const double g[2] = {-gp, gp};
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const double xi = g[i], eta = g[j];
      const double N[4] = {0.25 * (1 - xi) * (1 - eta), 0.25 * (1 + xi) * (1 - eta),
                           0.25 * (1 + xi) * (1 + eta), 0.25 * (1 - xi) * (1 + eta)};
      const double dxi[4] = {-0.25 * (1 - eta), 0.25 * (1 - eta), 0.25 * (1 + eta), -0.25 * (1 + eta)};
      const double deta[4] = {-0.25 * (1 - xi), -0.25 * (1 + xi), 0.25 * (1 + xi), 0.25 * (1 - xi)};
      double a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, p = 0.0;
      for (int k = 0; k < 4; ++k) {
        for (int d = 0; d < 3; ++d) {
          a[d] += P(face[k], d) * dxi[k];
          b[d] += P(face[k], d) * deta[k];
        }
        p += N[k] * pr[k];
      }
      const double av[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                            a[0] * b[1] - a[1] * b[0]};
      for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) f24[3 * face[k] + d] += (-1.0 * p) * av[d] * N[k];
    }
  return ORC_OK;
}

}  // extern "C"
