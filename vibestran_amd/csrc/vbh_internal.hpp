// vibestran_amd/csrc/vbh_internal.hpp -- shared declarations of libvbh.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/vbh.h"

namespace vbh {

// ---- errors ----------------------------------------------------------------
struct Error {
  int code;
  std::string msg;
};
void set_error(const std::string& msg);
[[noreturn]] void throw_error(int code, const std::string& msg);

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return VBH_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host out of memory");
    return VBH_ERR_HIP;
  } catch (const std::exception& e) {
    set_error(e.what());
    return VBH_ERR_ARG;
  }
}

#define VBH_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      ::vbh::throw_error(VBH_ERR_HIP, std::string("HIP error '") +                 \
                                          hipGetErrorString(e_) + "' at " #call); \
  } while (0)

// ---- device buffers ----------------------------------------------------------
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void resize(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    VBH_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    resize(count);
    if (count) VBH_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
};

// ---- grid reduction scratch ---------------------------------------------------
constexpr int kBlock = 256;            // threads per workgroup (4 wave64)
constexpr int kMaxGrid = 2048;         // grid-stride cap: bounds the partial sums
constexpr int kMaxRed = 4;             // values per grid reduction

// Device-side PCG scalars. One instance per right-hand side.
struct PcgScalars {
  double rz, rz_new, pAp, rr, alpha, beta, rel, norm_b;
  double rr_refresh;   // recursive rr at the last true-residual refresh (mixed)
  double rel_true;     // last FP64 true relative residual (mixed)
  double rel_prev;     // previous true relative residual (stall detection)
  int iter, done, flags, refreshes;
  int stalls, pad0, pad1, pad2;
};

enum : int {
  kFlagBreakdown = 1,    // p.Ap <= 0 (cuda_pcg_backend_impl.hpp:691-696)
  kFlagConverged = 2,
  kFlagMaxIter = 4,
  kFlagRz0 = 8,          // r.z <= 0 at start (impl:1859-1862)
  kFlagZeroRhs = 16,     // ||F|| < 1e-300 (impl:1822-1830)
  kFlagRefresh = 32,     // mixed: paused, FP64 true residual requested
  kFlagStall = 64        // mixed: true residual stopped improving
};

struct PcgParams {
  double tol;
  int max_iters;
  double refresh_delta;  // mixed: refresh when rr < delta^2 * rr_refresh
};

// Reduction workspace: per-value partials and one arrival counter per use.
struct RedWs {
  double* partials;   // kMaxRed * kMaxGrid
  unsigned* counter;  // one counter
};

// ---- operator formats ---------------------------------------------------------
// Scalar CSR (full symmetric) on the device.
struct CsrDev {
  int n = 0, nnz = 0;
  const int* rp = nullptr;
  const int* ci = nullptr;
  const double* v64 = nullptr;
  const float* v32 = nullptr;
};
// Block CSR with B x B blocks. Values of block row I (nb blocks from offset o)
// are stored entry-major: val[B*B*o + e*nb + k] is entry e (row-major in the
// block) of the row's k-th block, so lanes that own consecutive blocks read
// consecutive addresses.
struct BsrDev {
  int B = 0, nbr = 0, nnzb = 0;
  const int* brp = nullptr;
  const int* bci = nullptr;
  const double* v64 = nullptr;
  const float* v32 = nullptr;
};

// ---- kernel launchers (vbh_kernels.hip) ------------------------------------------
int grid_for(long long work_items, int items_per_block);

// y = A x ; optionally dot partial of x.y reduced into red (finalize via mode).
enum class SpmvEpilogue : int { None = 0, PcgAlpha = 1, TrueResidual = 2 };

template <typename T>
void launch_spmv_csr(const CsrDev& A, const T* x, T* y, hipStream_t s);
template <typename T>
void launch_spmv_bsr(const BsrDev& A, const T* x, T* y, hipStream_t s);

// PCG building blocks (vector type T: double or float; dots in double).
template <typename T>
void launch_pcg_init(int n, const double* F, const T* diag, double* u64, T* r, T* z, T* p,
                     PcgScalars* st, RedWs ws, hipStream_t s);
template <typename T>
void launch_pcg_spmv_alpha(int format, const CsrDev& A, const BsrDev& Ab, int n, const T* p,
                           T* Ap, PcgScalars* st, RedWs ws, hipStream_t s);
template <typename T>
void launch_pcg_update(int n, int mixed, const T* diag, const T* p, const T* Ap, double* u64,
                       T* r, T* z, PcgScalars* st, PcgParams prm, RedWs ws, hipStream_t s);
template <typename T>
void launch_pcg_pupdate(int n, const T* z, T* p, PcgScalars* st, hipStream_t s);
// Mixed refresh: r64 = F - K64 u64 (FP64), decide convergence, reload r32/z32
// and continue the recurrence with the refreshed residual.
void launch_refresh_true_residual(int format, const CsrDev& A, const BsrDev& Ab, int n,
                                  const double* F, const double* u64, double* r64, PcgScalars* st,
                                  PcgParams prm, RedWs ws, hipStream_t s);
void launch_refresh_reload(int n, const double* r64, const float* diag, float* r, float* z,
                           float* p, PcgScalars* st, PcgParams prm, RedWs ws, hipStream_t s);

// Plain elementwise helpers.
void launch_cast_f64_to_f32(int n, const double* a, float* b, hipStream_t s);
void launch_fill_f64(long long n, double* a, double v, hipStream_t s);

}  // namespace vbh
