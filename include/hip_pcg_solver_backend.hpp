// include/hip_pcg_solver_backend.hpp -- the hip-pcg SolverBackend as a
// maintainer adds it to the reference (as include/solver/hip_pcg_solver_backend.hpp).
//
// A subclass of vibestran::SolverBackend (reference
// include/solver/solver_backend.hpp:20-51) over vbh::HipPcgSolver
// (include/vbh_backend.hpp, the C ABI of vbh.h). Compiled only inside the
// reference's include tree: it needs "solver/solver_backend.hpp" and
// "core/exceptions.hpp". Registered next to the other GPU backends in the
// CLI's backend switch (reference src/main.cpp:276-343).
#pragma once

#include <optional>
#include <string_view>
#include <utility>
#include <vector>

#include "core/exceptions.hpp"
#include "solver/solver_backend.hpp"
#include "vbh_backend.hpp"

namespace vibestran {

class HipPcgSolverBackend final : public SolverBackend {
 public:
  // CudaPCGSolverBackend::try_create (reference
  // include/solver/cuda_pcg_solver_backend.hpp:45-47): nullopt without a GPU.
  [[nodiscard]] static std::optional<HipPcgSolverBackend> try_create(double tolerance = 0.0,
                                                                     int max_iters = 0,
                                                                     bool mixed = false) noexcept {
    auto s = vbh::HipPcgSolver::try_create(tolerance, max_iters,
                                           mixed ? VBH_PREC_MIXED : VBH_PREC_FP64);
    if (!s) return std::nullopt;
    return HipPcgSolverBackend(std::move(*s));
  }

  [[nodiscard]] std::vector<double> solve(
      const SparseMatrixBuilder::CsrData& K, const std::vector<double>& F,
      const FactorRatioCheckPolicy* /*factor_ratio_policy*/ = nullptr) override {
    try {
      return s_.solve(K, F);  // K may be Lower (the assembly's output)
    } catch (const vbh::SolverError& e) {
      throw SolverError(e.what());  // the message texts of cuda-pcg's SolverErrors
    }
  }
  [[nodiscard]] std::string_view name() const noexcept override { return s_.name(); }
  [[nodiscard]] int last_iteration_count() const noexcept override {
    return s_.last_iteration_count();
  }
  [[nodiscard]] double last_estimated_error() const noexcept override {
    return s_.last_estimated_error();
  }
  [[nodiscard]] std::string_view device_name() const noexcept { return s_.device_name(); }

 private:
  explicit HipPcgSolverBackend(vbh::HipPcgSolver s) noexcept : s_(std::move(s)) {}
  vbh::HipPcgSolver s_;
};

}  // namespace vibestran
