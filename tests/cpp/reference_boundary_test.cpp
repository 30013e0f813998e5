// tests/cpp/reference_boundary_test.cpp -- include/hip_pcg_solver_backend.hpp
// compiled against the reference's own headers (solver_backend.hpp,
// core/sparse_matrix.hpp, core/exceptions.hpp, used where they lie) and run
// through a SolverBackend pointer as LinearStaticSolver::solve_subcase calls
// it (reference src/solver/linear_static.cpp:266-285). Exit 77 when no GPU is
// usable (try_create -> nullopt), 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <memory>
#include <optional>
#include <type_traits>
#include <vector>

#include "hip_pcg_solver_backend.hpp"

using vibestran::HipPcgSolverBackend;
using vibestran::SolverBackend;
using vibestran::SparseMatrixBuilder;
using vibestran::SymmetryStorage;

static_assert(std::is_base_of_v<SolverBackend, HipPcgSolverBackend>);
static_assert(std::is_nothrow_move_constructible_v<HipPcgSolverBackend>);
static_assert(!std::is_copy_constructible_v<HipPcgSolverBackend>);
static_assert(std::is_same_v<decltype(HipPcgSolverBackend::try_create()),
                             std::optional<HipPcgSolverBackend>>);

#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::printf("FAILED %s at line %d\n", #c, __LINE__);                \
      return 1;                                                           \
    }                                                                     \
  } while (0)

// Lower CSR of the 1D Laplacian-like tridiagonal [-1 d -1], as build_csr
// returns it for an assembled model (SymmetryStorage::Lower).
static SparseMatrixBuilder::CsrData tridiag_lower(int n, double d) {
  SparseMatrixBuilder::CsrData K;
  K.n = n;
  K.row_ptr.push_back(0);
  for (int i = 0; i < n; ++i) {
    if (i > 0) {
      K.col_ind.push_back(i - 1);
      K.values.push_back(-1.0);
    }
    K.col_ind.push_back(i);
    K.values.push_back(d);
    K.row_ptr.push_back(static_cast<int>(K.col_ind.size()));
  }
  K.nnz = static_cast<int>(K.values.size());
  K.symmetry = SymmetryStorage::Lower;
  return K;
}

int main() {
  auto made = HipPcgSolverBackend::try_create(1e-10, 5000);
  if (!made) {
    std::printf("no GPU: try_create returned nullopt\n");
    return 77;
  }
  std::unique_ptr<SolverBackend> backend = std::make_unique<HipPcgSolverBackend>(std::move(*made));
  CHECK(!backend->name().empty());
  CHECK(!backend->requires_full_symmetric_csr());

  const int n = 500;
  const auto K = tridiag_lower(n, 2.5);
  std::vector<double> F(n);
  for (int i = 0; i < n; ++i) F[i] = std::sin(0.01 * i) + 1.0;
  const std::vector<double> u = backend->solve(K, F);  // default policy argument
  CHECK(static_cast<int>(u.size()) == n);
  double rr = 0.0, ff = 0.0;
  for (int i = 0; i < n; ++i) {
    const double Ku = 2.5 * u[i] - (i > 0 ? u[i - 1] : 0.0) - (i + 1 < n ? u[i + 1] : 0.0);
    rr += (F[i] - Ku) * (F[i] - Ku);
    ff += F[i] * F[i];
  }
  CHECK(std::sqrt(rr / ff) < 1e-9);
  CHECK(backend->last_iteration_count() > 0);
  CHECK(backend->last_estimated_error() >= 0.0 && backend->last_estimated_error() < 1e-10);

  // a zero diagonal: vibestran::SolverError with the cuda-pcg message shape
  auto Kz = tridiag_lower(8, 2.0);
  Kz.values[static_cast<size_t>(Kz.row_ptr[3] + 1)] = 0.0;  // diagonal of row 3
  bool threw = false;
  try {
    (void)backend->solve(Kz, std::vector<double>(8, 1.0));
  } catch (const vibestran::SolverError& e) {
    threw = std::string(e.what()).find("zero diagonal at row 3") != std::string::npos;
  }
  CHECK(threw);
  std::printf("OK reference boundary: %s, %d iterations\n", std::string(backend->name()).c_str(),
              backend->last_iteration_count());
  return 0;
}
