// include/vbh_backend.hpp -- C++20 host mirror of vibestran's solver plugin
// (reference include/solver/solver_backend.hpp:20-51 and the CUDA PCG
// backend's factory/diagnostics, include/solver/cuda_pcg_solver_backend.hpp:29-76)
// over the C ABI of vbh.h. Header-only; link with libvbh.so.
//
// HipPcgSolver::solve accepts any CSR type shaped like
// SparseMatrixBuilder::CsrData (reference include/core/sparse_matrix.hpp:63-87):
// members row_ptr, col_ind, values (std::vector), n, nnz and
// stores_lower_triangle_only(). Failures throw vbh::SolverError carrying the
// library's message (the reference's SolverError text shapes).
#pragma once

#include <algorithm>
#include <optional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

#include "vbh.h"

namespace vbh {

// vibestran::SolverError (reference include/core/exceptions.hpp:14-17).
class SolverError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

class HipPcgSolver {
 public:
  // CudaPCGSolverBackend::try_create: nullopt when no GPU is usable;
  // tolerance <= 0 selects 1e-8, max_iters <= 0 selects 10000.
  [[nodiscard]] static std::optional<HipPcgSolver> try_create(double tolerance = 0.0,
                                                              int max_iters = 0,
                                                              int precision = VBH_PREC_FP64,
                                                              int device = 0) noexcept {
    vbh_ctx* h = nullptr;
    if (vbh_ctx_create(device, tolerance, max_iters, &h) != VBH_OK || !h) return std::nullopt;
    if (vbh_set_precision(h, precision) != VBH_OK) {
      vbh_ctx_destroy(h);
      return std::nullopt;
    }
    return HipPcgSolver(h, precision);
  }

  HipPcgSolver(HipPcgSolver&& o) noexcept { *this = std::move(o); }
  HipPcgSolver& operator=(HipPcgSolver&& o) noexcept {
    if (this != &o) {
      release();
      h_ = std::exchange(o.h_, nullptr);
      precision_ = o.precision_;
      key_ = o.key_;
      last_iters_ = o.last_iters_;
      last_rel_ = o.last_rel_;
    }
    return *this;
  }
  HipPcgSolver(const HipPcgSolver&) = delete;
  HipPcgSolver& operator=(const HipPcgSolver&) = delete;
  ~HipPcgSolver() { release(); }

  // SolverBackend::solve: K u = F; K may store the lower triangle only.
  template <class Csr>
  [[nodiscard]] std::vector<double> solve(const Csr& K, const std::vector<double>& F) {
    if (K.n == 0) throw SolverError("HIP PCG: stiffness matrix is empty -- no free DOFs");
    if (static_cast<int>(F.size()) != K.n)
      throw SolverError("HIP PCG: force vector size " + std::to_string(F.size()) +
                        " != matrix size " + std::to_string(K.n));
    upload(K);
    std::vector<double> u(F.size());
    int it = 0;
    double rel = 0.0;
    const int rc = vbh_solve(h_, F.data(), u.data(), &it, &rel);
    last_iters_ = it;
    last_rel_ = rel;
    check(rc);
    return u;
  }

  // k load cases sharing K: F and the result are column-major n x k.
  template <class Csr>
  [[nodiscard]] std::vector<double> solve_block(const Csr& K, const std::vector<double>& F,
                                                int k) {
    if (k <= 0 || static_cast<long long>(F.size()) != static_cast<long long>(K.n) * k)
      throw SolverError("HIP PCG: block force array must be n x k");
    upload(K);
    std::vector<double> U(F.size());
    std::vector<int> its(static_cast<size_t>(k));
    std::vector<double> rels(static_cast<size_t>(k));
    const int rc = vbh_solve_block(h_, k, F.data(), U.data(), its.data(), rels.data());
    last_iters_ = 0;
    last_rel_ = 0.0;
    for (int j = 0; j < k; ++j) {
      last_iters_ = std::max(last_iters_, its[static_cast<size_t>(j)]);
      last_rel_ = std::max(last_rel_, rels[static_cast<size_t>(j)]);
    }
    check(rc);
    return U;
  }

  [[nodiscard]] std::string_view name() const noexcept {
    return precision_ == VBH_PREC_MIXED ? "HIP PCG + Jacobi float32 / float64 refinement (GPU)"
                                        : "HIP PCG + Jacobi float64 (GPU)";
  }
  [[nodiscard]] int last_iteration_count() const noexcept { return last_iters_; }
  [[nodiscard]] double last_relative_residual() const noexcept { return last_rel_; }
  [[nodiscard]] double last_estimated_error() const noexcept { return last_rel_; }
  [[nodiscard]] bool requires_full_symmetric_csr() const noexcept { return false; }
  [[nodiscard]] std::string_view device_name() const noexcept { return vbh_device_name(h_); }
  [[nodiscard]] vbh_ctx* handle() noexcept { return h_; }
  // Forget the uploaded operator (after changing a CSR's values in place).
  void invalidate() noexcept { key_ = Key{}; }

 private:
  HipPcgSolver(vbh_ctx* h, int precision) noexcept : h_(h), precision_(precision) {}

  void release() noexcept {
    if (h_) vbh_ctx_destroy(h_);
    h_ = nullptr;
  }

  static void check(int rc) {
    if (rc == VBH_OK) return;
    throw SolverError(vbh_last_error());
  }

  // The operator is re-uploaded only when the CSR object or its arrays change.
  template <class Csr>
  void upload(const Csr& K) {
    const Key key{&K, K.values.data(), K.n, K.nnz};
    if (key == key_) return;
    check(vbh_set_csr(h_, K.n, K.nnz, K.row_ptr.data(), K.col_ind.data(), K.values.data(),
                      K.stores_lower_triangle_only() ? 1 : 0));
    key_ = key;
  }

  struct Key {
    const void* obj = nullptr;
    const double* vals = nullptr;
    int n = -1, nnz = -1;
    bool operator==(const Key& o) const {
      return obj == o.obj && vals == o.vals && n == o.n && nnz == o.nnz;
    }
  };

  vbh_ctx* h_ = nullptr;
  int precision_ = VBH_PREC_FP64;
  Key key_;
  int last_iters_ = 0;
  double last_rel_ = 0.0;
};

}  // namespace vbh
