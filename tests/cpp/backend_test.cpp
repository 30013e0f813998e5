// tests/cpp/backend_test.cpp -- the C++ host mirror (include/vbh_backend.hpp)
// driven the way the reference's LinearStatic driver drives a SolverBackend
// (reference src/solver/linear_static.cpp:266-285), against closed-form
// answers. Exit 0 and "OK" on success; 77 when no GPU is visible.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "vbh_backend.hpp"

namespace {

enum class SymmetryStorage { Full, Lower };

// Shape of SparseMatrixBuilder::CsrData (reference include/core/sparse_matrix.hpp:63-87).
struct CsrData {
  std::vector<int> row_ptr, col_ind;
  std::vector<double> values;
  int n = 0, nnz = 0;
  SymmetryStorage symmetry = SymmetryStorage::Full;
  bool stores_lower_triangle_only() const noexcept { return symmetry == SymmetryStorage::Lower; }
};

// Lower storage of tridiag(-1, d, -1).
CsrData tridiag(int n, double d) {
  CsrData K;
  K.n = n;
  K.symmetry = SymmetryStorage::Lower;
  K.row_ptr.push_back(0);
  for (int i = 0; i < n; ++i) {
    if (i) {
      K.col_ind.push_back(i - 1);
      K.values.push_back(-1.0);
    }
    K.col_ind.push_back(i);
    K.values.push_back(d);
    K.row_ptr.push_back(static_cast<int>(K.col_ind.size()));
  }
  K.nnz = static_cast<int>(K.col_ind.size());
  return K;
}

// Thomas algorithm for the same matrix.
std::vector<double> thomas(int n, double d, const std::vector<double>& f) {
  std::vector<double> c(n), g(n), x(n);
  c[0] = -1.0 / d;
  g[0] = f[0] / d;
  for (int i = 1; i < n; ++i) {
    const double m = d + c[i - 1];
    c[i] = -1.0 / m;
    g[i] = (f[i] + g[i - 1]) / m;
  }
  x[n - 1] = g[n - 1];
  for (int i = n - 2; i >= 0; --i) x[i] = g[i] - c[i] * x[i + 1];
  return x;
}

int failures = 0;
void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAIL: %s\n", what);
    ++failures;
  }
}

}  // namespace

int main() {
  auto solver = vbh::HipPcgSolver::try_create(1e-12, 5000);
  if (!solver) {
    std::printf("no GPU visible\n");
    return 77;
  }
  const int n = 200;
  const CsrData K = tridiag(n, 2.5);
  std::vector<double> F(n);
  for (int i = 0; i < n; ++i) F[static_cast<size_t>(i)] = std::sin(0.1 * i) + 0.5;
  const std::vector<double> x = thomas(n, 2.5, F);
  const std::vector<double> u = solver->solve(K, F);
  double err = 0.0, xm = 0.0;
  for (int i = 0; i < n; ++i) {
    err = std::fmax(err, std::fabs(u[static_cast<size_t>(i)] - x[static_cast<size_t>(i)]));
    xm = std::fmax(xm, std::fabs(x[static_cast<size_t>(i)]));
  }
  expect(err <= 1e-10 * xm, "tridiagonal solve matches the Thomas algorithm");
  expect(solver->last_iteration_count() > 0 && solver->last_estimated_error() < 1e-12,
         "diagnostics after solve");
  expect(solver->name() == "HIP PCG + Jacobi float64 (GPU)", "backend name");

  // three load cases in one block solve
  std::vector<double> F3(3 * static_cast<size_t>(n));
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < n; ++i) F3[static_cast<size_t>(j * n + i)] = (j + 1) * F[static_cast<size_t>(i)];
  const std::vector<double> U3 = solver->solve_block(K, F3, 3);
  double berr = 0.0;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < n; ++i)
      berr = std::fmax(berr, std::fabs(U3[static_cast<size_t>(j * n + i)] - (j + 1) * x[static_cast<size_t>(i)]));
  expect(berr <= 3e-10 * xm, "block solve of three scaled load cases");

  // non-convergence and singular operator raise SolverError
  auto tight = vbh::HipPcgSolver::try_create(1e-12, 3);
  bool threw = false;
  try {
    (void)tight->solve(tridiag(n, 2.01), F);
  } catch (const vbh::SolverError& e) {
    threw = std::string(e.what()).find("did not converge after 3 iterations") != std::string::npos;
  }
  expect(threw, "max-iteration SolverError");
  CsrData Z = tridiag(4, 2.0);
  Z.values[0] = 0.0;
  threw = false;
  try {
    (void)solver->solve(Z, std::vector<double>(4, 1.0));
  } catch (const vbh::SolverError& e) {
    threw = std::string(e.what()).find("zero diagonal at row 0") != std::string::npos;
  }
  expect(threw, "zero-diagonal SolverError");

  // mixed precision reaches the FP64 tolerance through refreshes
  auto mixed = vbh::HipPcgSolver::try_create(1e-10, 5000, VBH_PREC_MIXED);
  const std::vector<double> um = mixed->solve(K, F);
  double merr = 0.0;
  for (int i = 0; i < n; ++i)
    merr = std::fmax(merr, std::fabs(um[static_cast<size_t>(i)] - x[static_cast<size_t>(i)]));
  expect(merr <= 1e-8 * xm, "mixed-precision solve");

  if (failures) return 1;
  std::printf("OK\n");
  return 0;
}
