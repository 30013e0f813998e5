// vibestran_amd/csrc/vbh_kernels.hip -- SpMV and Jacobi-PCG kernels for gfx950.
//
// Design (wave64, CDNA4):
//  * SpMV is HBM-bound (0.17 flop/B in FP64), so every kernel streams the
//    matrix exactly once with coalesced loads and keeps x in L2/MALL.
//  * Scalar CSR: a group of G lanes per row, shuffle reduction.
//  * BSR (B x B blocks): 32 lanes per block row for B=3 (27 blocks per node in
//    a hex mesh), 16 for B=6; each lane owns one block, values are stored
//    entry-major per block row so the lanes of a row load consecutive words.
//  * PCG dot products are fused into the producing kernel and reduced in a
//    fixed order (per-block partials + last-arriving block), so results are
//    deterministic for a given grid. All scalars (alpha, beta, convergence)
//    stay on the device; kernels exit at entry once `done` is set, which lets
//    the host enqueue many iterations (hipGraph) and poll only occasionally.
// Reference semantics (not code) followed: Jacobi-PCG of
// cuda_pcg_backend_impl.hpp:1795-1999, kernels 629-751.
#include "vbh_internal.hpp"

#include <cmath>

namespace vbh {

int grid_for(long long work_items, int items_per_block) {
  long long g = (work_items + items_per_block - 1) / items_per_block;
  if (g < 1) g = 1;
  if (g > kMaxGrid) g = kMaxGrid;
  return static_cast<int>(g);
}

namespace {

template <typename T>
__device__ __forceinline__ const T* vals(const CsrDev& A);
template <>
__device__ __forceinline__ const double* vals<double>(const CsrDev& A) { return A.v64; }
template <>
__device__ __forceinline__ const float* vals<float>(const CsrDev& A) { return A.v32; }
template <typename T>
__device__ __forceinline__ const T* bvals(const BsrDev& A);
template <>
__device__ __forceinline__ const double* bvals<double>(const BsrDev& A) { return A.v64; }
template <>
__device__ __forceinline__ const float* bvals<float>(const BsrDev& A) { return A.v32; }

// ---- deterministic reductions ---------------------------------------------------
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] += __shfl_down(v[i], off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) lds[w * NV + i] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0.0;
      for (int k = 0; k < nw; ++k) s += lds[k * NV + i];
      v[i] = s;
    }
  }
  __syncthreads();
}

// Sums v over the whole grid. Returns true in exactly one block (the last to
// arrive); there thread 0 holds the totals. The partials are summed in block
// index order, independent of arrival order.
template <int NV>
__device__ bool grid_sum(double (&v)[NV], RedWs ws, double* lds, int* s_last) {
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) ws.partials[i * kMaxGrid + blockIdx.x] = v[i];
    __threadfence();
    const unsigned prev = atomicAdd(ws.counter, 1u);
    *s_last = (prev == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!*s_last) return false;
  __threadfence();
  double t[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) t[i] = 0.0;
  for (int b = threadIdx.x; b < static_cast<int>(gridDim.x); b += blockDim.x)
#pragma unroll
    for (int i = 0; i < NV; ++i)
      t[i] += __builtin_nontemporal_load(&ws.partials[i * kMaxGrid + b]);
  block_sum<NV>(t, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = t[i];
    atomicExch(ws.counter, 0u);
  }
  return true;
}

// ---- row kernels ---------------------------------------------------------------
// Sum of row r of a scalar CSR times x, computed by G lanes; result in all lanes.
template <typename T, int G>
__device__ __forceinline__ T csr_row(const CsrDev& A, const T* __restrict__ val,
                                     const T* __restrict__ x, int r, int lg) {
  T s = T(0);
  const int e = A.rp[r + 1];
  for (int k = A.rp[r] + lg; k < e; k += G) s += val[k] * x[A.ci[k]];
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
  return s;
}

// B partial sums of block row I (G lanes, one block per lane); reduced in all lanes.
template <typename T, int B, int G>
__device__ __forceinline__ void bsr_row(const BsrDev& A, const T* __restrict__ val,
                                        const T* __restrict__ x, int I, int lg, T (&acc)[B]) {
#pragma unroll
  for (int r = 0; r < B; ++r) acc[r] = T(0);
  const int o = A.brp[I];
  const int nb = A.brp[I + 1] - o;
  const T* vb = val + static_cast<size_t>(B * B) * o;
  for (int k = lg; k < nb; k += G) {
    const int J = A.bci[o + k];
    T xj[B];
#pragma unroll
    for (int c = 0; c < B; ++c) xj[c] = x[B * J + c];
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int c = 0; c < B; ++c) acc[r] += vb[(r * B + c) * nb + k] * xj[c];
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
    for (int r = 0; r < B; ++r) acc[r] += __shfl_xor(acc[r], off, G);
}

template <int B>
struct BsrLanes {
  static constexpr int G = (B == 3) ? 32 : 16;
};

// ---- plain SpMV -------------------------------------------------------------------
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_spmv_csr(CsrDev A, const T* __restrict__ x,
                                                     T* __restrict__ y) {
  const T* val = vals<T>(A);
  const int lg = threadIdx.x % G;
  const int stride = gridDim.x * (blockDim.x / G);
  for (int r = (blockIdx.x * blockDim.x + threadIdx.x) / G; r < A.n; r += stride) {
    const T s = csr_row<T, G>(A, val, x, r, lg);
    if (lg == 0) y[r] = s;
  }
}

template <typename T, int B>
__global__ __launch_bounds__(kBlock) void k_spmv_bsr(BsrDev A, const T* __restrict__ x,
                                                     T* __restrict__ y) {
  constexpr int G = BsrLanes<B>::G;
  const T* val = bvals<T>(A);
  const int lg = threadIdx.x % G;
  const int stride = gridDim.x * (blockDim.x / G);
  for (int I = (blockIdx.x * blockDim.x + threadIdx.x) / G; I < A.nbr; I += stride) {
    T acc[B];
    bsr_row<T, B, G>(A, val, x, I, lg, acc);
    if (lg == 0)
#pragma unroll
      for (int r = 0; r < B; ++r) y[B * I + r] = acc[r];
  }
}

// ---- PCG: SpMV + p.Ap + alpha -------------------------------------------------------
__device__ __forceinline__ void finalize_alpha(PcgScalars* st, double pAp) {
  st->pAp = pAp;
  if (!(pAp > 0.0)) {  // breakdown (cuda_pcg_backend_impl.hpp:691-696)
    st->alpha = 0.0;
    st->flags |= kFlagBreakdown;
    st->rel = 1e300;
    st->done = 1;
    return;
  }
  st->alpha = st->rz / pAp;
}

template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_pcg_spmv_csr(CsrDev A, const T* __restrict__ p,
                                                         T* __restrict__ Ap, PcgScalars* st,
                                                         RedWs ws) {
  __shared__ double lds[4 * kMaxRed];
  __shared__ int s_last;
  if (st->done) return;
  const T* val = vals<T>(A);
  const int lg = threadIdx.x % G;
  const int stride = gridDim.x * (blockDim.x / G);
  double d[1] = {0.0};
  for (int r = (blockIdx.x * blockDim.x + threadIdx.x) / G; r < A.n; r += stride) {
    const T s = csr_row<T, G>(A, val, p, r, lg);
    if (lg == 0) {
      Ap[r] = s;
      d[0] += static_cast<double>(p[r]) * static_cast<double>(s);
    }
  }
  if (grid_sum<1>(d, ws, lds, &s_last) && threadIdx.x == 0) finalize_alpha(st, d[0]);
}

template <typename T, int B>
__global__ __launch_bounds__(kBlock) void k_pcg_spmv_bsr(BsrDev A, const T* __restrict__ p,
                                                         T* __restrict__ Ap, PcgScalars* st,
                                                         RedWs ws) {
  constexpr int G = BsrLanes<B>::G;
  __shared__ double lds[4 * kMaxRed];
  __shared__ int s_last;
  if (st->done) return;
  const T* val = bvals<T>(A);
  const int lg = threadIdx.x % G;
  const int stride = gridDim.x * (blockDim.x / G);
  double d[1] = {0.0};
  for (int I = (blockIdx.x * blockDim.x + threadIdx.x) / G; I < A.nbr; I += stride) {
    T acc[B];
    bsr_row<T, B, G>(A, val, p, I, lg, acc);
    if (lg == 0) {
#pragma unroll
      for (int r = 0; r < B; ++r) {
        Ap[B * I + r] = acc[r];
        d[0] += static_cast<double>(p[B * I + r]) * static_cast<double>(acc[r]);
      }
    }
  }
  if (grid_sum<1>(d, ws, lds, &s_last) && threadIdx.x == 0) finalize_alpha(st, d[0]);
}

// ---- PCG: init ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_init(int n, const double* __restrict__ F,
                                                     const T* __restrict__ diag,
                                                     double* __restrict__ u64, T* __restrict__ r,
                                                     T* __restrict__ z, T* __restrict__ p,
                                                     PcgScalars* st, RedWs ws) {
  __shared__ double lds[4 * kMaxRed];
  __shared__ int s_last;
  double v[3] = {0.0, 0.0, 0.0};  // ||F||^2, r.z, r.r (precision T)
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double f = F[i];
    const T ri = static_cast<T>(f);
    const T zi = ri / diag[i];
    u64[i] = 0.0;
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
    v[0] += f * f;
    v[1] += static_cast<double>(ri) * static_cast<double>(zi);
    v[2] += static_cast<double>(ri) * static_cast<double>(ri);
  }
  if (grid_sum<3>(v, ws, lds, &s_last) && threadIdx.x == 0) {
    st->iter = 0;
    st->flags = 0;
    st->done = 0;
    st->refreshes = 0;
    st->stalls = 0;
    st->alpha = st->beta = 0.0;
    st->pAp = 0.0;
    st->rel = 1.0;
    st->rel_true = 1.0;
    st->rel_prev = 1e300;
    const double norm_b = sqrt(fabs(v[0]));
    st->norm_b = norm_b;
    st->rz = v[1];
    st->rr = v[2];
    st->rr_refresh = v[2];
    if (norm_b < 1e-300) {  // impl:1822-1830: zero displacement, 0 iterations
      st->flags |= kFlagZeroRhs;
      st->rel = 0.0;
      st->done = 1;
    } else if (v[1] <= 0.0) {  // impl:1859-1862
      st->flags |= kFlagRz0;
      st->done = 1;
    }
  }
}

// ---- PCG: u += alpha p, r -= alpha Ap, z = r/diag, dots, beta -----------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int n, int mixed, const T* __restrict__ diag,
                                                       const T* __restrict__ p,
                                                       const T* __restrict__ Ap,
                                                       double* __restrict__ u64,
                                                       T* __restrict__ r, T* __restrict__ z,
                                                       PcgScalars* st, PcgParams prm, RedWs ws) {
  __shared__ double lds[4 * kMaxRed];
  __shared__ int s_last;
  if (st->done) return;
  const double alpha = st->alpha;
  const T a = static_cast<T>(alpha);
  double v[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const T pi = p[i];
    u64[i] += alpha * static_cast<double>(pi);
    const T ri = r[i] - a * Ap[i];
    const T zi = ri / diag[i];
    r[i] = ri;
    z[i] = zi;
    v[0] += static_cast<double>(ri) * static_cast<double>(zi);
    v[1] += static_cast<double>(ri) * static_cast<double>(ri);
  }
  if (grid_sum<2>(v, ws, lds, &s_last) && threadIdx.x == 0) {
    const double rz_new = v[0], rr = v[1];
    st->rz_new = rz_new;
    st->rr = rr;
    const double rel = sqrt(fabs(rr)) / st->norm_b;
    st->rel = rel;
    st->iter += 1;
    if (!mixed) {
      if (rel < prm.tol) {  // finalize_iteration_kernel, impl:704-738
        st->flags |= kFlagConverged;
        st->beta = 0.0;
        st->rz = rz_new;
        st->done = 1;
        return;
      }
    } else if (rel < prm.tol || rr < prm.refresh_delta * prm.refresh_delta * st->rr_refresh) {
      // Pause for an FP64 true-residual refresh (the recurrence resumes with
      // the refreshed residual; beta is computed then).
      st->flags |= kFlagRefresh;
      st->done = 1;
      return;
    }
    st->beta = rz_new / st->rz;
    st->rz = rz_new;
    if (st->iter >= prm.max_iters) {
      st->flags |= kFlagMaxIter;
      st->done = 1;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_pupdate(int n, const T* __restrict__ z,
                                                        T* __restrict__ p, PcgScalars* st) {
  if (st->done) return;
  const T b = static_cast<T>(st->beta);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = z[i] + b * p[i];
}

// ---- mixed precision refresh --------------------------------------------------------
// r64 = F - K u64 in FP64 with the true-residual norm; decides convergence.
__device__ __forceinline__ void finalize_true_residual(PcgScalars* st, double rr, PcgParams prm) {
  const double rel = sqrt(fabs(rr)) / st->norm_b;
  st->rel_true = rel;
  st->refreshes += 1;
  if (rel < prm.tol) {
    st->rel = rel;
    st->flags = (st->flags & ~kFlagRefresh) | kFlagConverged;
    return;  // done stays set
  }
  // Stall: the true residual no longer follows the recursive one.
  if (rel > 0.9 * st->rel_prev) st->stalls += 1; else st->stalls = 0;
  st->rel_prev = rel;
  if (st->stalls >= 4) {
    st->rel = rel;
    st->flags = (st->flags & ~kFlagRefresh) | kFlagStall;
  }
}

template <int G>
__global__ __launch_bounds__(kBlock) void k_true_residual_csr(CsrDev A, const double* __restrict__ F,
                                                              const double* __restrict__ u,
                                                              double* __restrict__ r64,
                                                              PcgScalars* st, PcgParams prm,
                                                              RedWs ws) {
  __shared__ double lds[4 * kMaxRed];
  __shared__ int s_last;
  if (!(st->flags & kFlagRefresh)) return;
  const int lg = threadIdx.x % G;
  const int stride = gridDim.x * (blockDim.x / G);
  double v[1] = {0.0};
  for (int r = (blockIdx.x * blockDim.x + threadIdx.x) / G; r < A.n; r += stride) {
    const double s = csr_row<double, G>(A, A.v64, u, r, lg);
    if (lg == 0) {
      const double ri = F[r] - s;
      r64[r] = ri;
      v[0] += ri * ri;
    }
  }
  if (grid_sum<1>(v, ws, lds, &s_last) && threadIdx.x == 0) finalize_true_residual(st, v[0], prm);
}

template <int B>
__global__ __launch_bounds__(kBlock) void k_true_residual_bsr(BsrDev A, const double* __restrict__ F,
                                                              const double* __restrict__ u,
                                                              double* __restrict__ r64,
                                                              PcgScalars* st, PcgParams prm,
                                                              RedWs ws) {
  constexpr int G = BsrLanes<B>::G;
  __shared__ double lds[4 * kMaxRed];
  __shared__ int s_last;
  if (!(st->flags & kFlagRefresh)) return;
  const int lg = threadIdx.x % G;
  const int stride = gridDim.x * (blockDim.x / G);
  double v[1] = {0.0};
  for (int I = (blockIdx.x * blockDim.x + threadIdx.x) / G; I < A.nbr; I += stride) {
    double acc[B];
    bsr_row<double, B, G>(A, A.v64, u, I, lg, acc);
    if (lg == 0)
#pragma unroll
      for (int r = 0; r < B; ++r) {
        const double ri = F[B * I + r] - acc[r];
        r64[B * I + r] = ri;
        v[0] += ri * ri;
      }
  }
  if (grid_sum<1>(v, ws, lds, &s_last) && threadIdx.x == 0) finalize_true_residual(st, v[0], prm);
}

// r32 = (float) r64, z = r32/diag, then beta from the refreshed r.z.
__global__ __launch_bounds__(kBlock) void k_refresh_reload(int n, const double* __restrict__ r64,
                                                           const float* __restrict__ diag,
                                                           float* __restrict__ r,
                                                           float* __restrict__ z, PcgScalars* st,
                                                           PcgParams prm, RedWs ws) {
  __shared__ double lds[4 * kMaxRed];
  __shared__ int s_last;
  if (!(st->flags & kFlagRefresh)) return;
  double v[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float ri = static_cast<float>(r64[i]);
    const float zi = ri / diag[i];
    r[i] = ri;
    z[i] = zi;
    v[0] += static_cast<double>(ri) * static_cast<double>(zi);
    v[1] += static_cast<double>(ri) * static_cast<double>(ri);
  }
  if (grid_sum<2>(v, ws, lds, &s_last) && threadIdx.x == 0) {
    st->rz_new = v[0];
    st->rr = v[1];
    st->rr_refresh = v[1];
    st->beta = v[0] / st->rz;
    st->rz = v[0];
    st->flags &= ~kFlagRefresh;
    if (st->iter >= prm.max_iters) {
      st->flags |= kFlagMaxIter;
      return;  // done stays set
    }
    st->done = 0;
  }
}

__global__ void k_cast_f64_f32(int n, const double* __restrict__ a, float* __restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    b[i] = static_cast<float>(a[i]);
}

__global__ void k_fill_f64(long long n, double* __restrict__ a, double v) {
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    a[i] = v;
}

constexpr int kCsrG = 8;  // lanes per scalar CSR row

}  // namespace

// ---- launchers ------------------------------------------------------------------------
template <typename T>
void launch_spmv_csr(const CsrDev& A, const T* x, T* y, hipStream_t s) {
  hipLaunchKernelGGL((k_spmv_csr<T, kCsrG>), dim3(grid_for(A.n, kBlock / kCsrG)), dim3(kBlock), 0,
                     s, A, x, y);
}

template <typename T>
void launch_spmv_bsr(const BsrDev& A, const T* x, T* y, hipStream_t s) {
  if (A.B == 3)
    hipLaunchKernelGGL((k_spmv_bsr<T, 3>), dim3(grid_for(A.nbr, kBlock / BsrLanes<3>::G)),
                       dim3(kBlock), 0, s, A, x, y);
  else
    hipLaunchKernelGGL((k_spmv_bsr<T, 6>), dim3(grid_for(A.nbr, kBlock / BsrLanes<6>::G)),
                       dim3(kBlock), 0, s, A, x, y);
}

template <typename T>
void launch_pcg_init(int n, const double* F, const T* diag, double* u64, T* r, T* z, T* p,
                     PcgScalars* st, RedWs ws, hipStream_t s) {
  hipLaunchKernelGGL((k_pcg_init<T>), dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, n, F,
                     diag, u64, r, z, p, st, ws);
}

template <typename T>
void launch_pcg_spmv_alpha(int format, const CsrDev& A, const BsrDev& Ab, int n, const T* p, T* Ap,
                           PcgScalars* st, RedWs ws, hipStream_t s) {
  (void)n;
  if (format == 0)
    hipLaunchKernelGGL((k_pcg_spmv_csr<T, kCsrG>), dim3(grid_for(A.n, kBlock / kCsrG)),
                       dim3(kBlock), 0, s, A, p, Ap, st, ws);
  else if (format == 3)
    hipLaunchKernelGGL((k_pcg_spmv_bsr<T, 3>), dim3(grid_for(Ab.nbr, kBlock / BsrLanes<3>::G)),
                       dim3(kBlock), 0, s, Ab, p, Ap, st, ws);
  else
    hipLaunchKernelGGL((k_pcg_spmv_bsr<T, 6>), dim3(grid_for(Ab.nbr, kBlock / BsrLanes<6>::G)),
                       dim3(kBlock), 0, s, Ab, p, Ap, st, ws);
}

template <typename T>
void launch_pcg_update(int n, int mixed, const T* diag, const T* p, const T* Ap, double* u64, T* r,
                       T* z, PcgScalars* st, PcgParams prm, RedWs ws, hipStream_t s) {
  hipLaunchKernelGGL((k_pcg_update<T>), dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, n, mixed,
                     diag, p, Ap, u64, r, z, st, prm, ws);
}

template <typename T>
void launch_pcg_pupdate(int n, const T* z, T* p, PcgScalars* st, hipStream_t s) {
  hipLaunchKernelGGL((k_pcg_pupdate<T>), dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, n, z, p,
                     st);
}

void launch_refresh_true_residual(int format, const CsrDev& A, const BsrDev& Ab, int n,
                                  const double* F, const double* u64, double* r64, PcgScalars* st,
                                  PcgParams prm, RedWs ws, hipStream_t s) {
  (void)n;
  if (format == 0)
    hipLaunchKernelGGL((k_true_residual_csr<kCsrG>), dim3(grid_for(A.n, kBlock / kCsrG)),
                       dim3(kBlock), 0, s, A, F, u64, r64, st, prm, ws);
  else if (format == 3)
    hipLaunchKernelGGL((k_true_residual_bsr<3>), dim3(grid_for(Ab.nbr, kBlock / BsrLanes<3>::G)),
                       dim3(kBlock), 0, s, Ab, F, u64, r64, st, prm, ws);
  else
    hipLaunchKernelGGL((k_true_residual_bsr<6>), dim3(grid_for(Ab.nbr, kBlock / BsrLanes<6>::G)),
                       dim3(kBlock), 0, s, Ab, F, u64, r64, st, prm, ws);
}

void launch_refresh_reload(int n, const double* r64, const float* diag, float* r, float* z,
                           float* p, PcgScalars* st, PcgParams prm, RedWs ws, hipStream_t s) {
  hipLaunchKernelGGL(k_refresh_reload, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, n, r64,
                     diag, r, z, st, prm, ws);
  // p = z + beta p runs only when the reload resumed the iteration.
  hipLaunchKernelGGL((k_pcg_pupdate<float>), dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, n,
                     z, p, st);
}

void launch_cast_f64_to_f32(int n, const double* a, float* b, hipStream_t s) {
  hipLaunchKernelGGL(k_cast_f64_f32, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, n, a, b);
}

void launch_fill_f64(long long n, double* a, double v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, n, a, v);
}

template void launch_spmv_csr<double>(const CsrDev&, const double*, double*, hipStream_t);
template void launch_spmv_csr<float>(const CsrDev&, const float*, float*, hipStream_t);
template void launch_spmv_bsr<double>(const BsrDev&, const double*, double*, hipStream_t);
template void launch_spmv_bsr<float>(const BsrDev&, const float*, float*, hipStream_t);
template void launch_pcg_init<double>(int, const double*, const double*, double*, double*, double*,
                                      double*, PcgScalars*, RedWs, hipStream_t);
template void launch_pcg_init<float>(int, const double*, const float*, double*, float*, float*,
                                     float*, PcgScalars*, RedWs, hipStream_t);
template void launch_pcg_spmv_alpha<double>(int, const CsrDev&, const BsrDev&, int, const double*,
                                            double*, PcgScalars*, RedWs, hipStream_t);
template void launch_pcg_spmv_alpha<float>(int, const CsrDev&, const BsrDev&, int, const float*,
                                           float*, PcgScalars*, RedWs, hipStream_t);
template void launch_pcg_update<double>(int, int, const double*, const double*, const double*,
                                        double*, double*, double*, PcgScalars*, PcgParams, RedWs,
                                        hipStream_t);
template void launch_pcg_update<float>(int, int, const float*, const float*, const float*, double*,
                                       float*, float*, PcgScalars*, PcgParams, RedWs, hipStream_t);
template void launch_pcg_pupdate<double>(int, const double*, double*, PcgScalars*, hipStream_t);
template void launch_pcg_pupdate<float>(int, const float*, float*, PcgScalars*, hipStream_t);

}  // namespace vbh
