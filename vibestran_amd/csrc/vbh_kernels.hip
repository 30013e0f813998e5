// vibestran_amd/csrc/vbh_kernels.hip -- SpMV and Jacobi-PCG kernels for gfx950.
//
// Design (wave64, CDNA4):
//  * SpMV is HBM-bound (0.17 flop/B in FP64), so every kernel streams the
//    matrix exactly once with coalesced loads and keeps x in L2/MALL.
//  * Scalar CSR: a group of G lanes per row, shuffle reduction.
//  * BSR (B x B blocks): 32 lanes per block row for B=3 (27 blocks per node in
//    a hex mesh), 16 for B=6; each lane owns one block, values are stored
//    entry-major per block row so the lanes of a row load consecutive words.
//  * PCG dot products are fused into the producing kernel as per-block
//    partial sums that the next kernel reduces in a fixed order, so results
//    are deterministic for a given grid. All scalars (alpha, beta,
//    convergence) stay on the device; kernels exit at entry once `done` is
//    set, which lets the host enqueue many iterations (hipGraph) and poll
//    only occasionally.
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

// ---- Jacobi-PCG, two kernels per iteration ------------------------------------------
// Every kernel that needs a global sum reduces the per-block partial sums its
// predecessor wrote (fixed order, every block redundantly), so no atomics or
// in-kernel fences are needed and the scalars are bit-identical in all blocks.
// Only block 0 writes the scalar state, and never a field that other blocks
// of the same kernel read.
//
//   init      r = F, u = 0, p_{-1} = 0; partials (r.z, r.r, F.F)       -> partB
//   cg_spmv   reduce partB -> rz_k, rr_k; convergence / pause decision;
//             p_k = r_k/diag + beta_k p_{k-1} formed in the gather;
//             Ap_k = K p_k; partial p_k.Ap_k                           -> partA
//   cg_update reduce partA -> alpha_k; u += alpha p_k; r -= alpha Ap_k;
//             partials (r.z, r.r)                                       -> partB
// Mixed precision pauses in cg_spmv when the recursive residual fell by
// refresh_delta since the last refresh (or below tol); the host then runs
//   refresh_residual  r64 = F - K64 u (FP64)                            -> partR
//   refresh_reload    reduce partR: converged / stalled / continue with
//                     r = (float) r64                                    -> partB
// and resumes: the next cg_spmv continues the recurrence (beta from the
// refreshed r.z), as in reliable-update CG.

template <int NV>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ part, int nprod,
                                                double (&out)[NV], double* lds) {
  double t[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) t[i] = 0.0;
  for (int b = threadIdx.x; b < nprod; b += blockDim.x)
#pragma unroll
    for (int i = 0; i < NV; ++i) t[i] += part[i * kMaxGrid + b];
  block_sum<NV>(t, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) lds[4 * kMaxRed + i] = t[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) out[i] = lds[4 * kMaxRed + i];
  __syncthreads();
}

template <int NV>
__device__ __forceinline__ void write_partials(double (&v)[NV], double* __restrict__ part,
                                               double* lds) {
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) part[i * kMaxGrid + blockIdx.x] = v[i];
}

constexpr int kLds = 5 * kMaxRed;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_init(int n, CgBufs<T> b, PcgScalars* st,
                                                    CgRed red) {
  __shared__ double lds[kLds];
  double v[3] = {0.0, 0.0, 0.0};  // r.z, r.r (precision T), F.F
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double f = b.F[i];
    const T ri = static_cast<T>(f);
    const T zi = ri / b.diag[i];
    b.u64[i] = 0.0;
    b.r[i] = ri;
    b.p1[i] = T(0);  // p_{-1}
    v[0] += static_cast<double>(ri) * static_cast<double>(zi);
    v[1] += static_cast<double>(ri) * static_cast<double>(ri);
    v[2] += f * f;
  }
  write_partials<3>(v, red.partB, lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->iter_done = 0;
    st->iter_cur = 0;
    st->flags = 0;
    st->done = 0;
    st->refreshes = 0;
    st->hist_stalls[0] = st->hist_stalls[1] = 0;
    st->hist_rel[0] = st->hist_rel[1] = 1e300;
    st->alpha = st->beta = st->pAp = 0.0;
    st->rel = 1.0;
    st->rel_true = 1.0;
    st->rz[0] = st->rz[1] = 0.0;
  }
}

// Prologue shared by the CSR and BSR variants of cg_spmv. Returns false if the
// kernel must exit; otherwise beta_k for the on-the-fly search direction.
__device__ __forceinline__ bool cg_spmv_prologue(PcgScalars* st, const CgRed& red,
                                                 const PcgParams& prm, int mixed, double* lds,
                                                 int* s_i, int& k, double& beta) {
  if (threadIdx.x == 0) {
    s_i[0] = st->done;
    s_i[1] = st->iter_done;
    s_i[2] = st->flags;
  }
  __syncthreads();
  if (s_i[0]) return false;
  k = s_i[1];
  const bool resumed = (s_i[2] & kFlagResumed) != 0;
  double v[3];
  reduce_partials<3>(red.partB, red.nB, v, lds);
  const double rz = v[0], rr = v[1];
  const bool lead = (blockIdx.x == 0 && threadIdx.x == 0);
  if (k == 0 && !resumed) {
    const double norm_b = sqrt(fabs(v[2]));
    if (lead) {
      st->norm_b = norm_b;
      st->rr_refresh = rr;
      st->rr = rr;
      st->rz[0] = rz;
      st->iter_cur = 0;
    }
    if (norm_b < 1e-300) {  // zero load: u = 0 after 0 iterations (impl:1822-1830)
      if (lead) {
        st->rel = 0.0;
        st->flags |= kFlagZeroRhs;
        st->done = 1;
      }
      return false;
    }
    if (rz <= 0.0) {  // impl:1859-1862
      if (lead) {
        st->rz[0] = rz;
        st->flags |= kFlagRz0;
        st->done = 1;
      }
      return false;
    }
    if (prm.max_iters <= 0) {
      if (lead) {
        st->flags |= kFlagMaxIter;
        st->done = 1;
      }
      return false;
    }
    beta = 0.0;
    return true;
  }
  const double rel = sqrt(fabs(rr)) / st->norm_b;
  int stop = 0;
  if (!mixed) {
    if (rel < prm.tol) stop = kFlagConverged;  // finalize_iteration, impl:704-738
  } else if (!resumed &&
             (rel < prm.tol || rr < prm.refresh_delta * prm.refresh_delta * st->rr_refresh)) {
    stop = kFlagRefresh;  // pause for an FP64 true-residual refresh
  }
  if (!stop && k >= prm.max_iters) stop = kFlagMaxIter;
  if (stop) {
    if (lead) {
      st->rel = rel;
      st->rr = rr;
      st->iter_cur = k;
      st->flags |= stop;
      st->done = 1;
    }
    return false;
  }
  beta = rz / st->rz[(k - 1) & 1];
  if (lead) {
    st->rz[k & 1] = rz;
    st->rr = rr;
    st->rel = rel;
    st->beta = beta;
    st->iter_cur = k;
    if (resumed) st->rr_refresh = rr;
  }
  return true;
}

template <typename T>
__device__ __forceinline__ T pdir(const CgBufs<T>& b, const T* __restrict__ pold, T beta, int j) {
  return b.r[j] / b.diag[j] + beta * pold[j];
}

template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_csr(CsrDev A, CgBufs<T> b, PcgScalars* st,
                                                        CgRed red, PcgParams prm, int mixed) {
  __shared__ double lds[kLds];
  __shared__ int s_i[4];
  int k;
  double betad;
  if (!cg_spmv_prologue(st, red, prm, mixed, lds, s_i, k, betad)) return;
  const T beta = static_cast<T>(betad);
  const T* pold = (k & 1) ? b.p0 : b.p1;
  T* pnew = (k & 1) ? b.p1 : b.p0;
  const T* val = vals<T>(A);
  const int lg = threadIdx.x % G;
  const int stride = gridDim.x * (blockDim.x / G);
  double d[1] = {0.0};
  for (int r = (blockIdx.x * blockDim.x + threadIdx.x) / G; r < A.n; r += stride) {
    T s = T(0);
    const int e = A.rp[r + 1];
    for (int q = A.rp[r] + lg; q < e; q += G) s += val[q] * pdir(b, pold, beta, A.ci[q]);
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
    if (lg == 0) {
      const T pr = pdir(b, pold, beta, r);
      pnew[r] = pr;
      b.Ap[r] = s;
      d[0] += static_cast<double>(pr) * static_cast<double>(s);
    }
  }
  write_partials<1>(d, red.partA, lds);
}

template <typename T, int B>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_bsr(BsrDev A, CgBufs<T> b, PcgScalars* st,
                                                        CgRed red, PcgParams prm, int mixed) {
  constexpr int G = BsrLanes<B>::G;
  __shared__ double lds[kLds];
  __shared__ int s_i[4];
  int k;
  double betad;
  if (!cg_spmv_prologue(st, red, prm, mixed, lds, s_i, k, betad)) return;
  const T beta = static_cast<T>(betad);
  const T* pold = (k & 1) ? b.p0 : b.p1;
  T* pnew = (k & 1) ? b.p1 : b.p0;
  const T* val = bvals<T>(A);
  const int lg = threadIdx.x % G;
  const int stride = gridDim.x * (blockDim.x / G);
  double d[1] = {0.0};
  for (int I = (blockIdx.x * blockDim.x + threadIdx.x) / G; I < A.nbr; I += stride) {
    T acc[B];
#pragma unroll
    for (int r = 0; r < B; ++r) acc[r] = T(0);
    const int o = A.brp[I];
    const int nb = A.brp[I + 1] - o;
    const T* vb = val + static_cast<size_t>(B * B) * o;
    for (int q = lg; q < nb; q += G) {
      const int J = A.bci[o + q];
      T xj[B];
#pragma unroll
      for (int c = 0; c < B; ++c) xj[c] = pdir(b, pold, beta, B * J + c);
#pragma unroll
      for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) acc[r] += vb[(r * B + c) * nb + q] * xj[c];
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
      for (int r = 0; r < B; ++r) acc[r] += __shfl_xor(acc[r], off, G);
    if (lg == 0) {
#pragma unroll
      for (int r = 0; r < B; ++r) {
        const T pr = pdir(b, pold, beta, B * I + r);
        pnew[B * I + r] = pr;
        b.Ap[B * I + r] = acc[r];
        d[0] += static_cast<double>(pr) * static_cast<double>(acc[r]);
      }
    }
  }
  write_partials<1>(d, red.partA, lds);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update(int n, CgBufs<T> b, PcgScalars* st,
                                                      CgRed red) {
  __shared__ double lds[kLds];
  __shared__ int s_i[2];
  if (threadIdx.x == 0) {
    s_i[0] = st->done;
    s_i[1] = st->iter_cur;
  }
  __syncthreads();
  if (s_i[0]) return;
  const int k = s_i[1];
  double v1[1];
  reduce_partials<1>(red.partA, red.nA, v1, lds);
  const double pAp = v1[0];
  const bool lead = (blockIdx.x == 0 && threadIdx.x == 0);
  if (!(pAp > 0.0)) {  // breakdown (cuda_pcg_backend_impl.hpp:691-696)
    if (lead) {
      st->pAp = pAp;
      st->alpha = 0.0;
      st->rel = 1e300;
      st->flags |= kFlagBreakdown;
      st->done = 1;
    }
    return;
  }
  const double alpha = st->rz[k & 1] / pAp;
  const T a = static_cast<T>(alpha);
  const T* p = (k & 1) ? b.p1 : b.p0;
  double v[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const T pi = p[i];
    b.u64[i] += alpha * static_cast<double>(pi);
    const T ri = b.r[i] - a * b.Ap[i];
    const T zi = ri / b.diag[i];
    b.r[i] = ri;
    v[0] += static_cast<double>(ri) * static_cast<double>(zi);
    v[1] += static_cast<double>(ri) * static_cast<double>(ri);
  }
  write_partials<2>(v, red.partB, lds);
  if (lead) {
    st->pAp = pAp;
    st->alpha = alpha;
    st->iter_done = k + 1;
    st->flags &= ~kFlagResumed;
  }
}

// ---- mixed precision refresh ----------------------------------------------------------
template <int G>
__global__ __launch_bounds__(kBlock) void k_refresh_residual_csr(CsrDev A, const double* __restrict__ F,
                                                                 const double* __restrict__ u,
                                                                 double* __restrict__ r64,
                                                                 PcgScalars* st, CgRed red) {
  __shared__ double lds[kLds];
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
  write_partials<1>(v, red.partR, lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->refreshes += 1;
}

template <int B>
__global__ __launch_bounds__(kBlock) void k_refresh_residual_bsr(BsrDev A, const double* __restrict__ F,
                                                                 const double* __restrict__ u,
                                                                 double* __restrict__ r64,
                                                                 PcgScalars* st, CgRed red) {
  constexpr int G = BsrLanes<B>::G;
  __shared__ double lds[kLds];
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
  write_partials<1>(v, red.partR, lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->refreshes += 1;
}

// Decide from the FP64 true residual; on "continue" reload r = (float) r64.
__global__ __launch_bounds__(kBlock) void k_refresh_reload(int n, const double* __restrict__ r64,
                                                           const float* __restrict__ diag,
                                                           float* __restrict__ r, PcgScalars* st,
                                                           CgRed red, PcgParams prm) {
  __shared__ double lds[kLds];
  double v1[1];
  reduce_partials<1>(red.partR, red.nR, v1, lds);
  const double rel = sqrt(fabs(v1[0])) / st->norm_b;
  const bool lead = (blockIdx.x == 0 && threadIdx.x == 0);
  // stall history is double-buffered by refresh count (refresh_residual
  // advanced the count), so no block reads what block 0 writes here
  const int rc = st->refreshes;
  const double rel_prev = st->hist_rel[(rc - 1) & 1];
  int stalls = st->hist_stalls[(rc - 1) & 1];
  int stop = 0;
  if (rel < prm.tol) {
    stop = kFlagConverged;
  } else {
    stalls = (rel > 0.9 * rel_prev) ? stalls + 1 : 0;  // true residual stopped following
    if (stalls >= 4) stop = kFlagStall;
  }
  if (lead) {
    st->rel_true = rel;
    st->hist_rel[rc & 1] = rel;
    st->hist_stalls[rc & 1] = stalls;
  }
  if (stop) {
    if (lead) {
      st->rel = rel;
      st->flags = (st->flags & ~kFlagRefresh) | stop;
    }
    return;  // done stays set
  }
  double v[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float ri = static_cast<float>(r64[i]);
    const float zi = ri / diag[i];
    r[i] = ri;
    v[0] += static_cast<double>(ri) * static_cast<double>(zi);
    v[1] += static_cast<double>(ri) * static_cast<double>(ri);
  }
  write_partials<3>(v, red.partB, lds);
  if (lead) {
    st->flags = (st->flags & ~kFlagRefresh) | kFlagResumed;
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
void launch_cg_init(int n, const CgBufs<T>& b, PcgScalars* st, const CgRed& red, hipStream_t s) {
  hipLaunchKernelGGL((k_cg_init<T>), dim3(red.nB), dim3(kBlock), 0, s, n, b, st, red);
}

template <typename T>
void launch_cg_spmv(int format, const CsrDev& A, const BsrDev& Ab, const CgBufs<T>& b,
                    PcgScalars* st, const CgRed& red, const PcgParams& prm, int mixed,
                    hipStream_t s) {
  if (format == 0)
    hipLaunchKernelGGL((k_cg_spmv_csr<T, kCsrG>), dim3(red.nA), dim3(kBlock), 0, s, A, b, st, red,
                       prm, mixed);
  else if (format == 3)
    hipLaunchKernelGGL((k_cg_spmv_bsr<T, 3>), dim3(red.nA), dim3(kBlock), 0, s, Ab, b, st, red, prm,
                       mixed);
  else
    hipLaunchKernelGGL((k_cg_spmv_bsr<T, 6>), dim3(red.nA), dim3(kBlock), 0, s, Ab, b, st, red, prm,
                       mixed);
}

template <typename T>
void launch_cg_update(int n, const CgBufs<T>& b, PcgScalars* st, const CgRed& red, hipStream_t s) {
  hipLaunchKernelGGL((k_cg_update<T>), dim3(red.nB), dim3(kBlock), 0, s, n, b, st, red);
}

void launch_refresh_residual(int format, const CsrDev& A, const BsrDev& Ab, const double* F,
                             const double* u64, double* r64, PcgScalars* st, const CgRed& red,
                             hipStream_t s) {
  if (format == 0)
    hipLaunchKernelGGL((k_refresh_residual_csr<kCsrG>), dim3(red.nR), dim3(kBlock), 0, s, A, F, u64,
                       r64, st, red);
  else if (format == 3)
    hipLaunchKernelGGL((k_refresh_residual_bsr<3>), dim3(red.nR), dim3(kBlock), 0, s, Ab, F, u64,
                       r64, st, red);
  else
    hipLaunchKernelGGL((k_refresh_residual_bsr<6>), dim3(red.nR), dim3(kBlock), 0, s, Ab, F, u64,
                       r64, st, red);
}

void launch_refresh_reload(int n, const double* r64, const float* diag, float* r, PcgScalars* st,
                           const CgRed& red, const PcgParams& prm, hipStream_t s) {
  hipLaunchKernelGGL(k_refresh_reload, dim3(red.nB), dim3(kBlock), 0, s, n, r64, diag, r, st, red,
                     prm);
}

int rows_per_block(int format) {
  return format == 0 ? kBlock / kCsrG : (format == 3 ? kBlock / BsrLanes<3>::G : kBlock / BsrLanes<6>::G);
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
template void launch_cg_init<double>(int, const CgBufs<double>&, PcgScalars*, const CgRed&, hipStream_t);
template void launch_cg_init<float>(int, const CgBufs<float>&, PcgScalars*, const CgRed&, hipStream_t);
template void launch_cg_spmv<double>(int, const CsrDev&, const BsrDev&, const CgBufs<double>&,
                                     PcgScalars*, const CgRed&, const PcgParams&, int, hipStream_t);
template void launch_cg_spmv<float>(int, const CsrDev&, const BsrDev&, const CgBufs<float>&,
                                    PcgScalars*, const CgRed&, const PcgParams&, int, hipStream_t);
template void launch_cg_update<double>(int, const CgBufs<double>&, PcgScalars*, const CgRed&, hipStream_t);
template void launch_cg_update<float>(int, const CgBufs<float>&, PcgScalars*, const CgRed&, hipStream_t);

}  // namespace vbh
