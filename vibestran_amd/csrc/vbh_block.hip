// vibestran_amd/csrc/vbh_block.hip -- block Jacobi-PCG over k load cases that
// share one stiffness operator (gfx950).
//
// The reference solves every SUBCASE separately (linear_static.cpp:202-320,
// one SolverBackend::solve per load vector). Here the k recurrences advance
// together so the operator is streamed from HBM once per iteration for all of
// them:
//   * vectors are row-major n_pad x ldk (ldk = k rounded up to a power of two,
//     padding columns stay zero and inactive), so one row of all load cases is
//     one contiguous run and lanes map to load cases;
//   * the SpMM reads each stored B x B block once per column tile: a group of
//     KT lanes (one per load case) owns one block row, gathers the B rows of
//     the block column for its KT cases and accumulates B outputs per lane.
//     With one group per wave the block row index is wave-uniform, so block
//     values and indices are fetched by scalar loads; with 64/KT groups per
//     wave every lane is busy but values go through vector loads. The choice,
//     the column tile KT and the grid are picked per (operator, k, precision)
//     by timing the candidates (autotune below);
//   * every recurrence keeps the reference's scalar semantics
//     (cuda_pcg_backend_impl.hpp:1795-1999: u0 = 0, r = F, z = r/diag, zero
//     load -> 0 iterations, r.z <= 0 -> error, p.Ap <= 0 -> breakdown,
//     converged when ||r||/||F|| < tol) independently per load case: a column
//     that converges or breaks down is frozen while the others continue;
//   * global sums are per-block partials reduced by a one-block-per-column
//     scalar kernel (fixed order, deterministic), which also owns that
//     column's state, so no kernel reads what another block of it writes;
//   * iterations are replayed from a captured hipGraph; the host polls the
//     column states between replays;
//   * when at most half of the columns are still running, the running ones
//     are moved into a layout half as wide (or narrower) and continue there:
//     a frozen column no longer costs its share of the gathers, the vector
//     streams and the stores of every later iteration. Each column's vectors
//     and scalar state move unchanged, so its iterates are the same; the
//     finished columns' solutions wait in the full-width buffers.
// Mixed precision (FP32 r, p, Ap, Jacobi and SpMM, FP64 solution update)
// pauses a column when its recursive residual fell by refresh_delta since its
// last refresh (or below tol); the host then recomputes r64 = F - K u in FP64
// for all columns and reloads the paused ones (reliable-update continuation),
// deciding convergence on the FP64 true residual.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "vbh_ctx.hpp"

namespace vbh {

// Per-load-case solver state, owned by block `column` of the scalar kernels.
struct BlkCol {
  double rz_old, rz, rr, rel, norm_b, alpha, beta, pAp, rr_refresh, rel_true, hist_rel;
  int iters, flags, state, refreshes, stalls, resumed;
};
enum : int { kColActive = 0, kColPaused = 1, kColDone = 2 };

// Block operator in array-of-blocks layout, blocks column-major: entry (r, c)
// of block q at v[q * B * B + c * B + r] (scalar CSR: B = 1, the CSR values).
struct BlkOp {
  int B = 0, nbr = 0;
  const int* brp = nullptr;
  const int* bci = nullptr;
  const double* v64 = nullptr;
  const float* v32 = nullptr;
  // FP64 diagonal node blocks corrected for row-sum preserving FP32 rounding
  // (row-major B*B per block row); the FP32 SpMM skips the FP32 diagonal
  // block and adds this block's product in FP64 after the row.
  const double* dblk = nullptr;
  // the single-RHS sliced ELL of the same operator (relative or absolute
  // layout; nslices = 0 when there is none), read by the sliced-ELL SpMM
  BsrDev sl;
};

struct BlkTune {
  int cw = 0, lg = 0, nrb = 0;  // columns per lane, lanes per group, grid rows
  int xcd = 0;                  // XCD-contiguous row ranges
  int stage = 0;                // block values: 0 direct loads, 1 staged through LDS,
                                // 2 broadcast across the group's lanes (lg = 4, 8, 16),
                                // 3 the sliced ELL, one lane per row (lg = 1, cw = 8, 4 or 2)
  double ms = 0.0;
};

struct BlockState {
  // operator copy in array-of-blocks layout
  long long op_version = -1;
  int B = 0;
  DBuf<double> aos64;
  DBuf<float> aos32;
  BlkOp op;
  // work space for one (ldk, n_pad)
  int ldk = 0, lg2 = 0, npad = 0, k = 0;
  long long work_version = -1;
  DBuf<double> F, u, r64, rd, pd, Apd;
  // mixed precision: r, p, Ap in FP32, and the solution's increments since the
  // last fold into u (FP32, folded once per batch of iterations)
  DBuf<float> rf, pf, Apf, duf;
  DBuf<double> partA, partB, partR;
  DBuf<double> coefA, coefB;
  DBuf<int> actD, actU, reload, ones;
  DBuf<BlkCol> cols;
  BlkCol* cols_host = nullptr;
  int nvec = 1;  // grid of the vector kernels (partials per column)
  // autotune cache for the current operator: (ldk, precision) -> config
  std::map<std::pair<int, int>, BlkTune> tuned;
  BlkTune cur;
  // full-width work space kept aside while compacted layouts run
  struct Wide {
    bool on = false;
    int ldk = 0, lg2 = 0, k = 0, nvec = 1;
    DBuf<double> F, u, r64, rd, pd, Apd, partA, partB, partR, coefA, coefB;
    DBuf<float> rf, pf, Apf, duf;
    DBuf<int> actD, actU, reload, ones;
    DBuf<BlkCol> cols;
    void reset() {  // after a solve that did not finish its compaction
      for (DBuf<double>* d : {&F, &u, &r64, &rd, &pd, &Apd, &partA, &partB, &partR, &coefA, &coefB})
        d->release();
      for (DBuf<float>* f : {&rf, &pf, &Apf, &duf}) f->release();
      for (DBuf<int>* i : {&actD, &actU, &reload, &ones}) i->release();
      cols.release();
      on = false;
    }
  } wide;
  // captured iteration graph
  hipGraphExec_t graph = nullptr;
  int g_prec = -1, g_ipg = 0, g_max_iters = 0, g_cw = 0, g_lg = 0, g_nrb = 0, g_xcd = 0, g_stage = 0;
  double g_tol = 0.0, g_delta = 0.0;
  ~BlockState() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (cols_host) (void)hipHostFree(cols_host);
  }
  void drop_graph() {
    if (graph) (void)hipGraphExecDestroy(graph);
    graph = nullptr;
  }
};

void destroy_block_state(BlockState* b) { delete b; }

namespace {

constexpr int kBT = 256;  // threads per block of every block-PCG kernel

// ---- block reductions -----------------------------------------------------------------
__device__ __forceinline__ double block_total(double v, double* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBT / 64; ++w) s += lds[w];
  __syncthreads();
  return s;
}

__device__ __forceinline__ bool any_active(const int* __restrict__ act, int ldk) {
  int a = 0;
  for (int j = 0; j < ldk; ++j) a |= act[j];
  return a != 0;
}

// ---- wide loads ---------------------------------------------------------------------------
template <typename T, int N>
struct alignas(sizeof(T) * N) Vec {
  T v[N];
};

template <typename T, int N>
__device__ __forceinline__ void vload(const T* __restrict__ p, T (&x)[N]) {
  const Vec<T, N> t = *reinterpret_cast<const Vec<T, N>*>(p);
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = t.v[i];
}

template <typename T, int N>
__device__ __forceinline__ void vstore(T* __restrict__ p, const T (&x)[N]) {
  Vec<T, N> t;
#pragma unroll
  for (int i = 0; i < N; ++i) t.v[i] = x[i];
  *reinterpret_cast<Vec<T, N>*>(p) = t;
}

// ---- vector kernels -------------------------------------------------------------------------
// Thread t handles V = 16 / sizeof(T) consecutive elements of one row, columns
// c0 = (t V) mod ldk .. c0 + V - 1; the grid stride is a multiple of ldk, so
// each thread keeps its columns. Element e = row * ldk + column, row = e >> lg2.
template <typename T>
constexpr int vwidth() {
  return 16 / static_cast<int>(sizeof(T));
}

// Per-column block sums of NV values: lds holds kBT * V doubles; thread t < ldk
// writes column t's sums to part[(i * ldk + t) * kMaxGrid + blockIdx.x].
template <int V, int NV>
__device__ __forceinline__ void column_partials(const double (&v)[NV][V], int ldk, double* lds,
                                                double* __restrict__ part) {
  const int per = ldk / V;  // threads sharing one set of columns repeat every `per`
#pragma unroll
  for (int i = 0; i < NV; ++i) {
#pragma unroll
    for (int w = 0; w < V; ++w) lds[threadIdx.x * V + w] = v[i][w];
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < ldk) {
      const int t = threadIdx.x, w = t % V;
      double s = 0.0;
      for (int th = t / V; th < kBT; th += per) s += lds[th * V + w];
      part[(static_cast<size_t>(i) * ldk + t) * kMaxGrid + blockIdx.x] = s;
    }
    __syncthreads();
  }
}

#define VBH_BLK_LOOP(e)                                                                      \
  for (long long e = (blockIdx.x * static_cast<long long>(kBT) + threadIdx.x) * V; e < m; \
       e += static_cast<long long>(gridDim.x) * kBT * V)

template <typename T>
__global__ __launch_bounds__(kBT) void k_blk_init(long long m, int ldk, int lg2,
                                                  const double* __restrict__ F,
                                                  const T* __restrict__ diag,
                                                  double* __restrict__ u, T* __restrict__ r,
                                                  T* __restrict__ p, double* __restrict__ part) {
  constexpr int V = vwidth<T>();
  __shared__ double lds[kBT * V];
  double acc[3][V] = {};
  VBH_BLK_LOOP(e) {
    const T d = diag[e >> lg2];
    double f[V], uz[V];
    T ri[V], zero[V];
    vload<double, V>(F + e, f);
#pragma unroll
    for (int w = 0; w < V; ++w) {
      ri[w] = static_cast<T>(f[w]);
      const T zi = ri[w] / d;
      acc[0][w] += static_cast<double>(ri[w]) * static_cast<double>(zi);
      acc[1][w] += static_cast<double>(ri[w]) * static_cast<double>(ri[w]);
      acc[2][w] += f[w] * f[w];
      uz[w] = 0.0;
      zero[w] = T(0);
    }
    vstore<double, V>(u + e, uz);
    vstore<T, V>(r + e, ri);
    vstore<T, V>(p + e, zero);
  }
  column_partials<V, 3>(acc, ldk, lds, part);
}

// p = z + beta p with z = r / diag, for the columns still running, after
// adding the previous iteration's alpha p to the solution of the columns it
// updated (actU): u (U = double, FP64) or the FP32 increments du (U = float,
// mixed precision; each step summed in FP64 and rounded once). p is read
// here anyway, so the update kernel need not read it again.
template <typename T, typename U>
__global__ __launch_bounds__(kBT) void k_blk_direction(long long m, int ldk, int lg2,
                                                       const T* __restrict__ diag,
                                                       const T* __restrict__ r, T* __restrict__ p,
                                                       const double* __restrict__ coefB,
                                                       const int* __restrict__ actD,
                                                       U* __restrict__ u,
                                                       const double* __restrict__ coefA,
                                                       const int* __restrict__ actU) {
  constexpr int V = vwidth<T>();
  if (!any_active(actD, ldk) && !any_active(actU, ldk)) return;
  const int c0 = (threadIdx.x * V) & (ldk - 1);
  bool on[V], up[V];
  T bt[V];
  double al[V];
  int any_on = 0, any_up = 0;
#pragma unroll
  for (int w = 0; w < V; ++w) {
    on[w] = actD[c0 + w] != 0;
    up[w] = actU[c0 + w] != 0;
    any_on |= on[w];
    any_up |= up[w];
    bt[w] = static_cast<T>(coefB[c0 + w]);
    al[w] = coefA[c0 + w];
  }
  if (!any_on && !any_up) return;
  VBH_BLK_LOOP(e) {
    T pv[V];
    vload<T, V>(p + e, pv);
    if (any_up) {
      U uv[V];
      vload<U, V>(u + e, uv);
#pragma unroll
      for (int w = 0; w < V; ++w)
        if (up[w]) uv[w] = static_cast<U>(static_cast<double>(uv[w]) + al[w] * static_cast<double>(pv[w]));
      vstore<U, V>(u + e, uv);
    }
    if (any_on) {
      const T d = diag[e >> lg2];
      T rv[V];
      vload<T, V>(r + e, rv);
#pragma unroll
      for (int w = 0; w < V; ++w) pv[w] = on[w] ? rv[w] / d + bt[w] * pv[w] : pv[w];
      vstore<T, V>(p + e, pv);
    }
  }
}

// r -= alpha Ap for the columns updated this iteration; partials r.z, r.r
// per column (the solution's alpha p is added by the next direction kernel,
// or by k_blk_settle at the end of the batch).
template <typename T>
__global__ __launch_bounds__(kBT) void k_blk_update(long long m, int ldk, int lg2,
                                                    const T* __restrict__ diag,
                                                    T* __restrict__ r, const T* __restrict__ Ap,
                                                    const double* __restrict__ coefA,
                                                    const int* __restrict__ actU,
                                                    double* __restrict__ part) {
  constexpr int V = vwidth<T>();
  __shared__ double lds[kBT * V];
  if (!any_active(actU, ldk)) return;
  const int c0 = (threadIdx.x * V) & (ldk - 1);
  bool on[V];
  T a[V];
#pragma unroll
  for (int w = 0; w < V; ++w) {
    on[w] = actU[c0 + w] != 0;
    a[w] = static_cast<T>(coefA[c0 + w]);
  }
  double acc[2][V] = {};
  VBH_BLK_LOOP(e) {
    const T d = diag[e >> lg2];
    T rv[V], qv[V];
    vload<T, V>(r + e, rv);
    vload<T, V>(Ap + e, qv);
#pragma unroll
    for (int w = 0; w < V; ++w) {
      if (on[w]) rv[w] = rv[w] - a[w] * qv[w];
      const T zi = rv[w] / d;
      acc[0][w] += static_cast<double>(rv[w]) * static_cast<double>(zi);
      acc[1][w] += static_cast<double>(rv[w]) * static_cast<double>(rv[w]);
    }
    vstore<T, V>(r + e, rv);
  }
  column_partials<V, 2>(acc, ldk, lds, part);
}

// End of a batch of iterations: u += du + alpha p for the columns whose last
// update is still pending (actU), du = 0 (du == nullptr: FP64, no
// increments), so that u is current wherever the host looks (refresh,
// compaction, the end of the solve); actU is cleared by the next launch.
// settle = 0 only clears du.
template <typename T>
__global__ __launch_bounds__(kBT) void k_blk_settle(long long m, int ldk, int settle,
                                                    double* __restrict__ u, float* __restrict__ du,
                                                    const T* __restrict__ p,
                                                    const double* __restrict__ coefA,
                                                    const int* __restrict__ actU) {
  constexpr int V = 4;
  const int c0 = (threadIdx.x * V) & (ldk - 1);
  bool up[V];
  double al[V];
  int any_up = 0;
#pragma unroll
  for (int w = 0; w < V; ++w) {
    up[w] = settle && actU[c0 + w] != 0;
    any_up |= up[w];
    al[w] = coefA[c0 + w];
  }
  for (long long e = (blockIdx.x * static_cast<long long>(kBT) + threadIdx.x) * V; e < m;
       e += static_cast<long long>(gridDim.x) * kBT * V) {
    float dv[V];
    if (du) {
      if (settle) vload<float, V>(du + e, dv);
      float zero[V];
#pragma unroll
      for (int w = 0; w < V; ++w) zero[w] = 0.0f;
      vstore<float, V>(du + e, zero);
    }
    if (!settle || (!du && !any_up)) continue;
    double uv[V];
    vload<double, V>(u + e, uv);
    T pv[V];
    if (any_up) vload<T, V>(p + e, pv);
#pragma unroll
    for (int w = 0; w < V; ++w) {
      if (du) uv[w] += static_cast<double>(dv[w]);
      if (up[w]) uv[w] += al[w] * static_cast<double>(pv[w]);
    }
    vstore<double, V>(u + e, uv);
  }
}

// Mixed refresh: reload r = (float) r64 in the columns the scalar kernel
// released; partials r.z, r.r per column.
__global__ __launch_bounds__(kBT) void k_blk_reload(long long m, int ldk, int lg2,
                                                    const float* __restrict__ diag,
                                                    const double* __restrict__ r64,
                                                    float* __restrict__ r,
                                                    const int* __restrict__ reload,
                                                    double* __restrict__ part) {
  constexpr int V = vwidth<float>();
  __shared__ double lds[kBT * V];
  const int c0 = (threadIdx.x * V) & (ldk - 1);
  bool on[V];
#pragma unroll
  for (int w = 0; w < V; ++w) on[w] = reload[c0 + w] != 0;
  double acc[2][V] = {};
  VBH_BLK_LOOP(e) {
    const float d = diag[e >> lg2];
    double wv[V];
    float rv[V];
    vload<double, V>(r64 + e, wv);
    vload<float, V>(r + e, rv);
#pragma unroll
    for (int w = 0; w < V; ++w) {
      if (on[w]) rv[w] = static_cast<float>(wv[w]);
      const float zi = rv[w] / d;
      acc[0][w] += static_cast<double>(rv[w]) * static_cast<double>(zi);
      acc[1][w] += static_cast<double>(rv[w]) * static_cast<double>(rv[w]);
    }
    vstore<float, V>(r + e, rv);
  }
  column_partials<V, 2>(acc, ldk, lds, part);
}
#undef VBH_BLK_LOOP

// ---- SpMM ---------------------------------------------------------------------------------
// MODE 0 (PCG):      Y = K X,      partial X.Y per column.
// MODE 1 (refresh):  Y = F - K X,  partial Y.Y per column (FP64).
// Lanes form groups of LG; a group owns one block row at a time and each of
// its lanes CW consecutive columns (one wide load per block row of X), so a
// wave covers 64 / LG block rows and a column tile of KT = LG * CW load cases.
// Grid: (nrb, ldk / KT), 4 waves per block.
template <typename T>
__device__ __forceinline__ const T* op_vals(const BlkOp& A);
template <>
__device__ __forceinline__ const double* op_vals<double>(const BlkOp& A) { return A.v64; }
template <>
__device__ __forceinline__ const float* op_vals<float>(const BlkOp& A) { return A.v32; }

// B consecutive values of a stored block (one block column; 16-byte loads
// where the layout allows).
template <typename T, int B>
__device__ __forceinline__ void block_row(const T* __restrict__ v, T (&a)[B]) {
  if constexpr (B == 6) {
    T t[2];
#pragma unroll
    for (int h = 0; h < 3; ++h) {
      vload<T, 2>(v + 2 * h, t);
      a[2 * h] = t[0];
      a[2 * h + 1] = t[1];
    }
  } else {
#pragma unroll
    for (int s = 0; s < B; ++s) a[s] = v[s];
  }
}

// xcd != 0: workgroup w runs on XCD w mod 8 (round-robin dispatch), so the
// block rows are split into 8 contiguous ranges, one per XCD, and each XCD's
// workgroups sweep their range together: rows that share block columns (the
// neighbours of a node) are then read through the same L2.
template <typename T, int B, int CW, int LG, int MODE>
__global__ __launch_bounds__(kBT) void k_blk_spmm(BlkOp A, int ldk, int xcd,
                                                  const T* __restrict__ X, T* __restrict__ Y,
                                                  const double* __restrict__ F,
                                                  const int* __restrict__ act,
                                                  double* __restrict__ part) {
  constexpr int KT = CW * LG, GR = 64 / LG;
  __shared__ double lds[kBT * CW];
  const int lane = threadIdx.x & 63;
  const int lg = lane % LG, g = lane / LG;
  const int c0 = blockIdx.y * KT + lg * CW;
  {  // a column tile with no column to compute leaves at once
    int on = 0;
#pragma unroll
    for (int w = 0; w < CW; ++w) on |= act[c0 + w];
    if (!__syncthreads_or(on)) return;
  }
  // FP32 products of 6x6 (shell) blocks are summed in FP64: their rows mix
  // membrane and bending terms of very different size, and FP32 sums cost
  // the row-sum preserving operator a large share of its iteration savings
  using Acc = typename std::conditional<(MODE == 0 && sizeof(T) == 4 && B == 6), double, T>::type;
  const T* __restrict__ val = op_vals<T>(A);
  const double* __restrict__ dblk = (MODE == 0 && sizeof(T) == 4) ? A.dblk : nullptr;
  const int wave = threadIdx.x >> 6;
  double dot[CW];
#pragma unroll
  for (int w = 0; w < CW; ++w) dot[w] = 0.0;
  int rb = blockIdx.x, nrbx = gridDim.x, row_lo = 0, row_hi = A.nbr;
  if (xcd) {
    const int x = blockIdx.x & 7, per = (A.nbr + 7) >> 3;
    rb = blockIdx.x >> 3;
    nrbx = gridDim.x >> 3;
    row_lo = min(A.nbr, x * per);
    row_hi = min(A.nbr, row_lo + per);
  }
  for (int I0 = row_lo + (rb * (kBT / 64) + wave) * GR; I0 < row_hi;
       I0 += nrbx * (kBT / 64) * GR) {
    const int I = I0 + g;
    if (I >= row_hi) continue;
    Acc acc[B][CW];
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int w = 0; w < CW; ++w) acc[r][w] = Acc(0);
    const int o = A.brp[I], e = A.brp[I + 1];
    for (int q = o; q < e; ++q) {
      const int J = A.bci[q];
      const T* __restrict__ v = val + static_cast<size_t>(q) * (B * B);
      const T* __restrict__ xj = X + static_cast<size_t>(B * J) * ldk + c0;
      if (dblk && J == I) continue;  // FP64 epilogue below
      // column s of the block (stored contiguously) times row s of X_J
#pragma unroll
      for (int s = 0; s < B; ++s) {
        T x[CW], a[B];
        vload<T, CW>(xj + static_cast<size_t>(s) * ldk, x);
        block_row<T, B>(v + s * B, a);
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
          for (int w = 0; w < CW; ++w) acc[r][w] += static_cast<Acc>(a[r]) * static_cast<Acc>(x[w]);
      }
    }
#pragma unroll
    for (int r = 0; r < B; ++r) {
      const size_t idx = static_cast<size_t>(B * I + r) * ldk + c0;
      if constexpr (MODE == 0) {
        T xo[CW], yo[CW];
        vload<T, CW>(X + idx, xo);
        for (int w = 0; w < CW; ++w) yo[w] = static_cast<T>(acc[r][w]);
        if (dblk) {  // FP64 corrected diagonal block: y = acc + D_II x_I in FP64
          double dsum[CW];
#pragma unroll
          for (int w = 0; w < CW; ++w) dsum[w] = static_cast<double>(acc[r][w]);
#pragma unroll
          for (int s = 0; s < B; ++s) {
            T xi[CW];
            vload<T, CW>(X + static_cast<size_t>(B * I + s) * ldk + c0, xi);
            const double dv = dblk[(static_cast<size_t>(I) * B + r) * B + s];
#pragma unroll
            for (int w = 0; w < CW; ++w) dsum[w] += dv * static_cast<double>(xi[w]);
          }
#pragma unroll
          for (int w = 0; w < CW; ++w) yo[w] = static_cast<T>(dsum[w]);
        }
        vstore<T, CW>(Y + idx, yo);
#pragma unroll
        for (int w = 0; w < CW; ++w)
          dot[w] += static_cast<double>(xo[w]) * static_cast<double>(yo[w]);
      } else {
        double f[CW];
        T y[CW];
        vload<double, CW>(F + idx, f);
#pragma unroll
        for (int w = 0; w < CW; ++w) {
          const double yy = f[w] - static_cast<double>(acc[r][w]);
          y[w] = static_cast<T>(yy);
          dot[w] += yy * yy;
        }
        vstore<T, CW>(Y + idx, y);
      }
    }
  }
  // per-column block sum: threads with equal lane % LG share columns
#pragma unroll
  for (int w = 0; w < CW; ++w) lds[threadIdx.x * CW + w] = dot[w];
  __syncthreads();
  if (static_cast<int>(threadIdx.x) < KT) {
    const int t = threadIdx.x, w = t % CW;
    double s = 0.0;
    for (int th = t / CW; th < kBT; th += LG) s += lds[th * CW + w];
    part[static_cast<size_t>(blockIdx.y * KT + t) * kMaxGrid + blockIdx.x] = s;
  }
}

// ---- scalar kernels: block j owns load case j -------------------------------------------
// mode 0: after init; 1: after an update; 2: after a refresh reload (only the
// reloaded columns). Reduces r.z, r.r (and F.F in mode 0) and decides.
__global__ __launch_bounds__(kBT) void k_blk_scal_beta(int mode, int ldk, int k, int mixed,
                                                       const double* __restrict__ part, int nvec,
                                                       BlkCol* __restrict__ cols,
                                                       double* __restrict__ coefB,
                                                       int* __restrict__ actD,
                                                       const int* __restrict__ reload,
                                                       PcgParams prm) {
  __shared__ double lds[kBT];
  const int j = blockIdx.x;
  BlkCol& cj = cols[j];
  if (mode == 2 && !reload[j]) return;  // untouched columns keep their coefficients
  if (mode == 1 && cj.state != kColActive) {
    if (threadIdx.x == 0) actD[j] = 0;
    return;
  }
  double v[3] = {0.0, 0.0, 0.0};
  const int nv = (mode == 0) ? 3 : 2;
  for (int i = 0; i < nv; ++i) {
    double t = 0.0;
    for (int b = threadIdx.x; b < nvec; b += kBT) t += part[(i * ldk + j) * kMaxGrid + b];
    v[i] = block_total(t, lds);
  }
  if (threadIdx.x != 0) return;
  const double rz = v[0], rr = v[1];
  int stop = 0;
  double beta = 0.0, rel;
  if (mode == 0) {
    cj.iters = 0;
    cj.flags = 0;
    cj.refreshes = 0;
    cj.stalls = 0;
    cj.resumed = 0;
    cj.hist_rel = 1e300;
    cj.alpha = cj.pAp = 0.0;
    cj.rel_true = 1.0;
    cj.rz_old = 0.0;
    cj.norm_b = sqrt(fabs(v[2]));
    rel = 1.0;
    if (j >= k) {  // padding column
      stop = kFlagZeroRhs;
      rel = 0.0;
    } else if (cj.norm_b < 1e-300) {  // zero load: u = 0 after 0 iterations
      stop = kFlagZeroRhs;
      rel = 0.0;
    } else if (rz <= 0.0) {
      stop = kFlagRz0;
    } else if (prm.max_iters <= 0) {
      stop = kFlagMaxIter;
    }
    if (!stop) cj.rr_refresh = rr;
  } else {
    rel = sqrt(fabs(rr)) / cj.norm_b;
    if (mode == 1) {
      if (!mixed) {
        if (rel < prm.tol) stop = kFlagConverged;
      } else if (rel < prm.tol || rr < prm.refresh_delta * prm.refresh_delta * cj.rr_refresh) {
        stop = kFlagRefresh;
      }
    }
    if (!stop && cj.iters >= prm.max_iters) stop = kFlagMaxIter;
    if (!stop) beta = rz / cj.rz_old;
    if (mode == 2) cj.rr_refresh = rr;
  }
  cj.rz = rz;
  cj.rr = rr;
  cj.rel = rel;
  if (stop == kFlagRefresh) {
    cj.state = kColPaused;
    actD[j] = 0;
    return;  // rz_old stays the predecessor's: beta after the reload uses it
  }
  if (stop) {
    cj.flags |= stop;
    cj.state = kColDone;
    actD[j] = 0;
    return;
  }
  cj.state = kColActive;
  cj.resumed = 0;
  cj.rz_old = rz;
  cj.beta = beta;
  coefB[j] = beta;
  actD[j] = 1;
}

// alpha = r.z / p.Ap for columns whose direction was formed this iteration.
__global__ __launch_bounds__(kBT) void k_blk_scal_alpha(int ldk, const double* __restrict__ part,
                                                        int nrb, BlkCol* __restrict__ cols,
                                                        double* __restrict__ coefA,
                                                        int* __restrict__ actU,
                                                        const int* __restrict__ actD) {
  __shared__ double lds[kBT];
  const int j = blockIdx.x;
  if (!actD[j]) {
    if (threadIdx.x == 0) {
      actU[j] = 0;
      coefA[j] = 0.0;
    }
    return;
  }
  double t = 0.0;
  for (int b = threadIdx.x; b < nrb; b += kBT) t += part[static_cast<size_t>(j) * kMaxGrid + b];
  const double pAp = block_total(t, lds);
  if (threadIdx.x != 0) return;
  BlkCol& cj = cols[j];
  cj.pAp = pAp;
  if (!(pAp > 0.0)) {  // breakdown (cuda_pcg_backend_impl.hpp:691-696)
    cj.flags |= kFlagBreakdown;
    cj.state = kColDone;
    cj.rel = 1e300;
    cj.alpha = 0.0;
    coefA[j] = 0.0;
    actU[j] = 0;
    return;
  }
  const double alpha = cj.rz_old / pAp;
  cj.alpha = alpha;
  cj.iters += 1;
  coefA[j] = alpha;
  actU[j] = 1;
}

// Mixed refresh decision for the paused columns from ||F - K u||.
__global__ __launch_bounds__(kBT) void k_blk_scal_refresh(int ldk, const double* __restrict__ part,
                                                          int nrb, BlkCol* __restrict__ cols,
                                                          int* __restrict__ reload, PcgParams prm) {
  __shared__ double lds[kBT];
  const int j = blockIdx.x;
  BlkCol& cj = cols[j];
  if (cj.state != kColPaused) {
    if (threadIdx.x == 0) reload[j] = 0;
    return;
  }
  double t = 0.0;
  for (int b = threadIdx.x; b < nrb; b += kBT) t += part[static_cast<size_t>(j) * kMaxGrid + b];
  const double rr = block_total(t, lds);
  if (threadIdx.x != 0) return;
  const double rel = sqrt(fabs(rr)) / cj.norm_b;
  cj.refreshes += 1;
  cj.rel_true = rel;
  int stop = 0;
  if (rel < prm.tol) {
    stop = kFlagConverged;
  } else {
    cj.stalls = (rel > 0.9 * cj.hist_rel) ? cj.stalls + 1 : 0;
    if (cj.stalls >= 4) stop = kFlagStall;
  }
  cj.hist_rel = rel;
  if (stop) {
    cj.rel = rel;
    cj.flags |= stop;
    cj.state = kColDone;
    reload[j] = 0;
    return;
  }
  cj.state = kColActive;
  cj.resumed = 1;
  reload[j] = 1;
}

// ---- layout helpers -------------------------------------------------------------------------
// Column-major n x k (leading dimension n) -> row-major npad x ldk, zero padded.
__global__ void k_to_rows(int n, int k, int npad, int ldk, const double* __restrict__ a,
                          double* __restrict__ b) {
  const long long m = static_cast<long long>(npad) * ldk;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < m;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int row = static_cast<int>(e / ldk), c = static_cast<int>(e % ldk);
    b[e] = (row < n && c < k) ? a[static_cast<size_t>(c) * n + row] : 0.0;
  }
}

__global__ void k_to_cols(int n, int k, int ldk, const double* __restrict__ b,
                          double* __restrict__ a) {
  const long long m = static_cast<long long>(n) * k;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < m;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int c = static_cast<int>(e / n), row = static_cast<int>(e % n);
    a[e] = b[static_cast<size_t>(row) * ldk + c];
  }
}

// Column moves between row-major layouts (npad rows): gather dst[row][j] =
// src[row][map[j]] (0 where map[j] < 0) into a narrower layout; scatter
// dst[row][map[j]] = src[row][j] where map[j] >= 0 back into a wider one.
template <typename T>
__global__ void k_blk_gather(long long npad, int ldk_src, int lg2_dst, const int* __restrict__ map,
                             const T* __restrict__ src, T* __restrict__ dst) {
  const long long m = npad << lg2_dst;
  const int mask = (1 << lg2_dst) - 1;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < m;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int j = map[static_cast<int>(e & mask)];
    dst[e] = j >= 0 ? src[(e >> lg2_dst) * ldk_src + j] : T(0);
  }
}

template <typename T>
__global__ void k_blk_scatter(long long npad, int lg2_src, int ldk_dst, const int* __restrict__ map,
                              const T* __restrict__ src, T* __restrict__ dst) {
  const long long m = npad << lg2_src;
  const int mask = (1 << lg2_src) - 1;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < m;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int j = map[static_cast<int>(e & mask)];
    if (j >= 0) dst[(e >> lg2_src) * ldk_dst + j] = src[e];
  }
}

// Entry-major BSR (vbh BsrDev) -> array-of-blocks, each block column-major
// (entry (r, c) at c * B + r), so the SpMM reads one block column per row of X.
template <int B>
__global__ void k_bsr_to_aos(int nbr, const int* __restrict__ brp, const double* __restrict__ em,
                             double* __restrict__ a64, float* __restrict__ a32) {
  for (int I = blockIdx.x * blockDim.x + threadIdx.x; I < nbr; I += gridDim.x * blockDim.x) {
    const int o = brp[I], nb = brp[I + 1] - o;
    for (int q = 0; q < nb; ++q)
#pragma unroll
      for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) {
          const double x = em[static_cast<size_t>(B * B) * o + static_cast<size_t>(r * B + c) * nb + q];
          a64[static_cast<size_t>(o + q) * (B * B) + c * B + r] = x;
          a32[static_cast<size_t>(o + q) * (B * B) + c * B + r] = static_cast<float>(x);
        }
  }
}

__global__ void k_fill_i32(int n, int* a, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}

// LDS-staged variant: at step t every group of the wave needs the t-th block
// of its row. The wave fetches those GR blocks with a few coalesced loads
// (each lane a distinct value) into its LDS slice, and each lane then reads
// its group's block column from LDS. A direct load of a block column would
// be broadcast to the LG lanes of a group but still occupy the address unit
// for all 64 lanes; through LDS the address unit sees B*B*GR values once.
__device__ __forceinline__ void wave_lds_sync() { __builtin_amdgcn_wave_barrier(); }

template <typename T, int B, int CW, int LG, int MODE>
__global__ __launch_bounds__(kBT) void k_blk_spmm_lds(BlkOp A, int ldk, int xcd,
                                                      const T* __restrict__ X, T* __restrict__ Y,
                                                      const double* __restrict__ F,
                                                      const int* __restrict__ act,
                                                      double* __restrict__ part) {
  constexpr int KT = CW * LG, GR = 64 / LG, NB2 = B * B, NV = GR * NB2;
  constexpr int NJ = (NV + 63) / 64;  // staged values per lane and step
  __shared__ T sA[kBT / 64][NV];
  __shared__ double lds[kBT * CW];
  const int lane = threadIdx.x & 63;
  const int lg = lane % LG, g = lane / LG;
  const int c0 = blockIdx.y * KT + lg * CW;
  {  // a column tile with no column to compute leaves at once
    int on = 0;
#pragma unroll
    for (int w = 0; w < CW; ++w) on |= act[c0 + w];
    if (!__syncthreads_or(on)) return;
  }
  // FP32 products of 6x6 (shell) blocks are summed in FP64: their rows mix
  // membrane and bending terms of very different size, and FP32 sums cost
  // the row-sum preserving operator a large share of its iteration savings
  using Acc = typename std::conditional<(MODE == 0 && sizeof(T) == 4 && B == 6), double, T>::type;
  const T* __restrict__ val = op_vals<T>(A);
  const double* __restrict__ dblk = (MODE == 0 && sizeof(T) == 4) ? A.dblk : nullptr;
  const int wave = threadIdx.x >> 6;
  T* __restrict__ stage = sA[wave];
  double dot[CW];
#pragma unroll
  for (int w = 0; w < CW; ++w) dot[w] = 0.0;
  int rb = blockIdx.x, nrbx = gridDim.x, row_lo = 0, row_hi = A.nbr;
  if (xcd) {
    const int x = blockIdx.x & 7, per = (A.nbr + 7) >> 3;
    rb = blockIdx.x >> 3;
    nrbx = gridDim.x >> 3;
    row_lo = min(A.nbr, x * per);
    row_hi = min(A.nbr, row_lo + per);
  }
  for (int I0 = row_lo + (rb * (kBT / 64) + wave) * GR; I0 < row_hi;
       I0 += nrbx * (kBT / 64) * GR) {
    const int I = I0 + g;
    const bool row_ok = I < row_hi;
    int o = 0, nb = 0;
    if (row_ok) {
      o = A.brp[I];
      nb = A.brp[I + 1] - o;
    }
    // owner group of each staged slot: its first block and block count
    int so[NJ], sn[NJ], se[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int f = lane + 64 * j;
      const int gj = (f < NV) ? f / NB2 : 0;
      se[j] = f - gj * NB2;
      // every lane takes part in the shuffles (a lane reading from an
      // inactive lane would get an undefined value)
      so[j] = __shfl(o, gj * LG, 64);
      const int nbj = __shfl(nb, gj * LG, 64);
      sn[j] = (f < NV) ? nbj : 0;
    }
    int nbmax = nb;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) nbmax = max(nbmax, __shfl_xor(nbmax, off, 64));
    Acc acc[B][CW];
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int w = 0; w < CW; ++w) acc[r][w] = Acc(0);
    for (int t = 0; t < nbmax; ++t) {
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int f = lane + 64 * j;
        if (f < NV)
          stage[f] = (t < sn[j]) ? val[static_cast<size_t>(so[j] + t) * NB2 + se[j]] : T(0);
      }
      wave_lds_sync();
      if (t < nb) {
        const int J = A.bci[o + t];
        const T* __restrict__ xj = X + static_cast<size_t>(B * J) * ldk + c0;
        const T* ag = stage + g * NB2;
        if (!(dblk && J == I)) {  // the diagonal block goes through the FP64 epilogue
#pragma unroll
          for (int s = 0; s < B; ++s) {
            T x[CW];
            vload<T, CW>(xj + static_cast<size_t>(s) * ldk, x);
#pragma unroll
            for (int r = 0; r < B; ++r)
#pragma unroll
              for (int w = 0; w < CW; ++w)
                acc[r][w] += static_cast<Acc>(ag[s * B + r]) * static_cast<Acc>(x[w]);
          }
        }
      }
      wave_lds_sync();
    }
    if (row_ok) {
#pragma unroll
      for (int r = 0; r < B; ++r) {
        const size_t idx = static_cast<size_t>(B * I + r) * ldk + c0;
        if constexpr (MODE == 0) {
          T xo[CW], yo[CW];
          vload<T, CW>(X + idx, xo);
          for (int w = 0; w < CW; ++w) yo[w] = static_cast<T>(acc[r][w]);
        if (dblk) {  // FP64 corrected diagonal block: y = acc + D_II x_I in FP64
          double dsum[CW];
#pragma unroll
          for (int w = 0; w < CW; ++w) dsum[w] = static_cast<double>(acc[r][w]);
#pragma unroll
          for (int s = 0; s < B; ++s) {
            T xi[CW];
            vload<T, CW>(X + static_cast<size_t>(B * I + s) * ldk + c0, xi);
            const double dv = dblk[(static_cast<size_t>(I) * B + r) * B + s];
#pragma unroll
            for (int w = 0; w < CW; ++w) dsum[w] += dv * static_cast<double>(xi[w]);
          }
#pragma unroll
          for (int w = 0; w < CW; ++w) yo[w] = static_cast<T>(dsum[w]);
        }
          vstore<T, CW>(Y + idx, yo);
#pragma unroll
          for (int w = 0; w < CW; ++w)
            dot[w] += static_cast<double>(xo[w]) * static_cast<double>(yo[w]);
        } else {
          double f[CW];
          T y[CW];
          vload<double, CW>(F + idx, f);
#pragma unroll
          for (int w = 0; w < CW; ++w) {
            const double yy = f[w] - static_cast<double>(acc[r][w]);
            y[w] = static_cast<T>(yy);
            dot[w] += yy * yy;
          }
          vstore<T, CW>(Y + idx, y);
        }
      }
    }
  }
#pragma unroll
  for (int w = 0; w < CW; ++w) lds[threadIdx.x * CW + w] = dot[w];
  __syncthreads();
  if (static_cast<int>(threadIdx.x) < KT) {
    const int t = threadIdx.x, w = t % CW;
    double sum = 0.0;
    for (int th = t / CW; th < kBT; th += LG) sum += lds[th * CW + w];
    part[static_cast<size_t>(blockIdx.y * KT + t) * kMaxGrid + blockIdx.x] = sum;
  }
}

// Register-broadcast variants (lanes per group LG = 4, 8 or 16): the LG lanes
// of a group load their block's B*B values with one coalesced load per LG
// values (lane l holds values l, l + LG, ...), and every product takes its
// block value from the owning lane with a broadcast inside the group, so the
// address unit sees each block value once:
//   LG = 16  DPP row_newbcast:n (gfx90a and later: lane n of each 16-lane row
//            to the whole row), folded by the compiler into the instruction
//            that consumes it (the FP32 -> FP64 conversion of the shell sums);
//            a plain DPP move for FP64 values (mov_dpp: every source lane is
//            in the row, so no old value; update_dpp with one costs a move
//            that initialises each destination);
//   LG = 4   DPP quad_perm [n,n,n,n], folded the same way;
//   LG = 8   ds_swizzle in bit-mask mode (lane (l & 0x18) | n of each 32),
//            which runs in the LDS unit without touching LDS memory.
constexpr int kDppRowNewBcast = 0x150;  // + n: row_newbcast:n

template <int LG, int N>
__device__ __forceinline__ int group_bcast_i32(int v) {
  static_assert(N >= 0 && N < LG, "source lane outside the group");
  if constexpr (LG == 16) {
    return __builtin_amdgcn_mov_dpp(v, kDppRowNewBcast + N, 0xF, 0xF, true);
  } else if constexpr (LG == 4) {
    return __builtin_amdgcn_mov_dpp(v, N | (N << 2) | (N << 4) | (N << 6), 0xF, 0xF, true);
  } else {
    static_assert(LG == 8, "group broadcast for 4, 8 or 16 lanes");
    return __builtin_amdgcn_ds_swizzle(v, 0x18 | (N << 5));
  }
}

template <int LG, int N>
__device__ __forceinline__ float group_bcast(float v) {
  return __int_as_float(group_bcast_i32<LG, N>(__float_as_int(v)));
}

template <int LG, int N>
__device__ __forceinline__ double group_bcast(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = group_bcast_i32<LG, N>(static_cast<int>(b));
  const int hi = group_bcast_i32<LG, N>(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

constexpr bool has_group_bcast(int lg) { return lg == 4 || lg == 8 || lg == 16; }

template <typename Fn, int... I>
__device__ __forceinline__ void static_for_seq(Fn&& fn, std::integer_sequence<int, I...>) {
  (fn(std::integral_constant<int, I>{}), ...);
}

template <int N, typename Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
  static_for_seq(fn, std::make_integer_sequence<int, N>{});
}

// values per lane of the FP64 diagonal block a group of LG lanes holds (only
// the FP32 plain product adds it)
template <typename T, int B, int LG, int MODE>
constexpr int dblk_lanes() {
  return (MODE == 0 && sizeof(T) == 4) ? (B * B + LG - 1) / LG : 0;
}

// No minimum of resident waves: capping the 6x6 FP32 shapes with two load
// cases per lane at 128 VGPRs (4 waves per SIMD) spills their accumulators.
template <typename T, int B, int CW, int LG, int MODE>
__global__ __launch_bounds__(kBT) void k_blk_spmm_bcast(
    BlkOp A, int ldk, int xcd, const T* __restrict__ X, T* __restrict__ Y,
    const double* __restrict__ F, const int* __restrict__ act, double* __restrict__ part) {
  constexpr int KT = CW * LG, GR = 64 / LG, NB2 = B * B, NL = (NB2 + LG - 1) / LG;
  __shared__ double lds[kBT * CW];
  const int lane = threadIdx.x & 63;
  const int lg = lane % LG, g = lane / LG;
  const int c0 = blockIdx.y * KT + lg * CW;
  {  // a column tile with no column to compute leaves at once
    int on = 0;
#pragma unroll
    for (int w = 0; w < CW; ++w) on |= act[c0 + w];
    if (!__syncthreads_or(on)) return;
  }
  // FP64 sums for the FP32 6x6 (shell) blocks, as in k_blk_spmm
  using Acc = typename std::conditional<(MODE == 0 && sizeof(T) == 4 && B == 6), double, T>::type;
  const T* __restrict__ val = op_vals<T>(A);
  const double* __restrict__ dblk = (MODE == 0 && sizeof(T) == 4) ? A.dblk : nullptr;
  const int wave = threadIdx.x >> 6;
  double dot[CW];
#pragma unroll
  for (int w = 0; w < CW; ++w) dot[w] = 0.0;
  int rb = blockIdx.x, nrbx = gridDim.x, row_lo = 0, row_hi = A.nbr;
  if (xcd) {
    const int x = blockIdx.x & 7, per = (A.nbr + 7) >> 3;
    rb = blockIdx.x >> 3;
    nrbx = gridDim.x >> 3;
    row_lo = min(A.nbr, x * per);
    row_hi = min(A.nbr, row_lo + per);
  }
  // every lane of a group follows the same block row and block sequence, so
  // the broadcast source lane is always active when its row is.
  // Rows are software-pipelined: the row pointers two rows ahead and the first
  // block columns of the next row load while this row is summed, so a row
  // starts with its block loads instead of two dependent index round trips;
  // where registers allow (kPreX), the row's own X rows load at its start too.
  constexpr bool kPreX = CW <= 2;
  const int step = nrbx * (kBT / 64) * GR;
  auto row_ptrs = [&](int I, int& ro, int& re) {
    ro = 0;
    re = 0;
    if (I < row_hi) {
      ro = A.brp[I];
      re = A.brp[I + 1];
    }
  };
  auto first_cols = [&](int ro, int re) { return (lg < min(LG, re - ro)) ? A.bci[ro + lg] : 0; };
  int I0 = row_lo + (rb * (kBT / 64) + wave) * GR;
  int o, e, o1, e1, o2, e2;
  row_ptrs(I0 + g, o, e);
  row_ptrs(I0 + step + g, o1, e1);
  int jl0 = first_cols(o, e), jl1 = 0;
  for (; I0 < row_hi; I0 += step, o = o1, e = e1, o1 = o2, e1 = e2, jl0 = jl1) {
    const int I = I0 + g;
    row_ptrs(I + 2 * step, o2, e2);
    jl1 = first_cols(o1, e1);
    if (I >= row_hi) continue;
    Acc acc[B][CW];
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int w = 0; w < CW; ++w) acc[r][w] = Acc(0);
    // the row's own X rows (for p.Ap, and the diagonal block)
    T xI[B][CW];
    if constexpr (MODE == 0 && kPreX) {
#pragma unroll
      for (int s = 0; s < B; ++s)
        vload<T, CW>(X + static_cast<size_t>(B * I + s) * ldk + c0, xI[s]);
    }
    // the row's block columns LG at a time (lane l holds the l-th); blocks in
    // pairs with two register sets, so the next block's values and X rows
    // load while the current one is summed (the diagonal block of an FP32
    // operator is neither loaded nor summed: FP64 epilogue below)
    T m0[NL], x0[B][CW], m1[NL], x1[B][CW];
    auto fetch = [&](T (&mv)[NL], T (&xv)[B][CW], int q, int J) {
      const T* __restrict__ v = val + static_cast<size_t>(q) * NB2;
#pragma unroll
      for (int j = 0; j < NL; ++j) {
        const int f = lg + LG * j;
        mv[j] = (f < NB2) ? v[f] : T(0);
      }
      const T* __restrict__ xj = X + static_cast<size_t>(B * J) * ldk + c0;
#pragma unroll
      for (int s = 0; s < B; ++s) vload<T, CW>(xj + static_cast<size_t>(s) * ldk, xv[s]);
    };
    auto sum = [&](const T (&mv)[NL], const T (&xv)[B][CW]) {
      // value i of the block is entry (r, s) = (i mod B, i / B) (column-major)
      static_for<NB2>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        constexpr int r = i % B, s = i / B;
        const Acc a = static_cast<Acc>(group_bcast<LG, i % LG>(mv[i / LG]));
#pragma unroll
        for (int w = 0; w < CW; ++w) acc[r][w] += a * static_cast<Acc>(xv[s][w]);
      });
    };
    for (int q0 = o; q0 < e; q0 += LG) {
      const int qn = min(LG, e - q0);
      const int jl = (q0 == o) ? jl0 : ((lg < qn) ? A.bci[q0 + lg] : 0);
      int J = __shfl(jl, 0, LG);
      bool use0 = !(dblk && J == I), use1 = false;
      if (use0) fetch(m0, x0, q0, J);
      for (int t = 0; t < qn; t += 2) {
        if (t + 1 < qn) {
          J = __shfl(jl, t + 1, LG);
          use1 = !(dblk && J == I);
          if (use1) fetch(m1, x1, q0 + t + 1, J);
        }
        if (use0) sum(m0, x0);
        if (t + 1 >= qn) break;
        if (t + 2 < qn) {
          J = __shfl(jl, t + 2, LG);
          use0 = !(dblk && J == I);
          if (use0) fetch(m0, x0, q0 + t + 2, J);
        }
        if (use1) sum(m1, x1);
      }
    }
    if constexpr (MODE == 0) {
      if constexpr (!kPreX) {
#pragma unroll
        for (int s = 0; s < B; ++s)
          vload<T, CW>(X + static_cast<size_t>(B * I + s) * ldk + c0, xI[s]);
      }
      // FP64 corrected diagonal block (row-major B*B per block row), held
      // like the block values: lane l of the group has entries l, l + 16, ...
      constexpr int ND = dblk_lanes<T, B, LG, MODE>();
      double dm[ND > 0 ? ND : 1];
      if constexpr (ND > 0) {
        if (dblk) {
#pragma unroll
          for (int j = 0; j < ND; ++j) {
            const int f = lg + LG * j;
            dm[j] = (f < NB2) ? dblk[static_cast<size_t>(I) * NB2 + f] : 0.0;
          }
        }
      }
      static_for<B>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        const size_t idx = static_cast<size_t>(B * I + r) * ldk + c0;
        T yo[CW];
#pragma unroll
        for (int w = 0; w < CW; ++w) yo[w] = static_cast<T>(acc[r][w]);
        if constexpr (ND > 0) {
          if (dblk) {  // y = acc + D_II x_I in FP64
            double dsum[CW];
#pragma unroll
            for (int w = 0; w < CW; ++w) dsum[w] = static_cast<double>(acc[r][w]);
            static_for<B>([&](auto sc) {
              constexpr int i = r * B + decltype(sc)::value;
              const double dv = group_bcast<LG, i % LG>(dm[i / LG]);
#pragma unroll
              for (int w = 0; w < CW; ++w)
                dsum[w] += dv * static_cast<double>(xI[decltype(sc)::value][w]);
            });
#pragma unroll
            for (int w = 0; w < CW; ++w) yo[w] = static_cast<T>(dsum[w]);
          }
        }
        vstore<T, CW>(Y + idx, yo);
#pragma unroll
        for (int w = 0; w < CW; ++w)
          dot[w] += static_cast<double>(xI[r][w]) * static_cast<double>(yo[w]);
      });
    } else {
#pragma unroll
      for (int r = 0; r < B; ++r) {
        const size_t idx = static_cast<size_t>(B * I + r) * ldk + c0;
        double f[CW];
        T y[CW];
        vload<double, CW>(F + idx, f);
#pragma unroll
        for (int w = 0; w < CW; ++w) {
          const double yy = f[w] - static_cast<double>(acc[r][w]);
          y[w] = static_cast<T>(yy);
          dot[w] += yy * yy;
        }
        vstore<T, CW>(Y + idx, y);
      }
    }
  }
#pragma unroll
  for (int w = 0; w < CW; ++w) lds[threadIdx.x * CW + w] = dot[w];
  __syncthreads();
  if (static_cast<int>(threadIdx.x) < KT) {
    const int t = threadIdx.x, w = t % CW;
    double sum = 0.0;
    for (int th = t / CW; th < kBT; th += LG) sum += lds[th * CW + w];
    part[static_cast<size_t>(blockIdx.y * KT + t) * kMaxGrid + blockIdx.x] = sum;
  }
}

// Sliced-ELL variant (value path 3): one lane per scalar row of a 64-row
// slice, streaming the single-RHS operator layout (BsrDev: relative slot
// offsets or absolute block columns), KC columns per lane gathered as whole
// X rows. Every block value is read once per column tile, by the one lane
// whose row it is on, with no cross-lane broadcast; the X gathers grow with
// KC, so it suits narrow column tiles. Same inputs, outputs and partial sums
// as the other shapes (column tile blockIdx.y of KC columns, part slot
// blockIdx.x).
template <typename T, int B, int KC, int MODE>
__global__ __launch_bounds__(kBT) void k_blk_spmm_sell(
    BlkOp A, int ldk, int xcd, const T* __restrict__ X, T* __restrict__ Y,
    const double* __restrict__ F, const int* __restrict__ act, double* __restrict__ part) {
  __shared__ double lds[kBT * KC];
  const int l = threadIdx.x & (kSliceRows - 1), wave = threadIdx.x >> 6;
  const int c0 = blockIdx.y * KC;
  {  // a column tile with no column to compute leaves at once
    int on = 0;
#pragma unroll
    for (int w = 0; w < KC; ++w) on |= act[c0 + w];
    if (!__syncthreads_or(on)) return;
  }
  using Acc = typename std::conditional<(MODE == 0 && sizeof(T) == 4 && B == 6), double, T>::type;
  constexpr int NPS = sell_nps(B);
  const BsrDev& S = A.sl;
  const bool rel = S.rk64 != nullptr;
  const int* __restrict__ kk = rel ? (sizeof(T) == 4 ? S.rk32 : S.rk64) : S.koff;
  const int* __restrict__ ro = sizeof(T) == 4 ? S.ro32 : S.ro64;
  const T* __restrict__ vals;
  if constexpr (sizeof(T) == 4) vals = rel ? S.rv32 : S.s32;
  else vals = rel ? S.rv64 : S.s64;
  const double* __restrict__ dblk = (MODE == 0 && sizeof(T) == 4) ? A.dblk : nullptr;
  const int nrows = S.nbr * B, last = S.nbr - 1;
  double dot[KC];
#pragma unroll
  for (int w = 0; w < KC; ++w) dot[w] = 0.0;
  // slices: XCD-contiguous ranges (xcd) or grid-stride, one per wave
  const int wpb = kBT / kSliceRows;
  int s0 = static_cast<int>(blockIdx.x) * wpb + wave, stride = static_cast<int>(gridDim.x) * wpb;
  int s_hi = S.nslices;
  if (xcd && gridDim.x >= 8) {
    const int x = blockIdx.x & 7, per = (S.nslices + 7) >> 3;
    const int lo = min(S.nslices, x * per);
    s0 = lo + static_cast<int>(blockIdx.x >> 3) * wpb + wave;
    stride = static_cast<int>(gridDim.x >> 3) * wpb;
    s_hi = min(S.nslices, lo + per);
  }
  for (int sc = s0; sc < s_hi; sc += stride) {
    const int sw = __builtin_amdgcn_readfirstlane(sc);
    const int k0 = kk[sw], km = kk[sw + 1] - k0;
    const int g = kSliceRows * sw + l;
    const int node = g / B, node0 = (kSliceRows * sw) / B;
    const T* __restrict__ v = vals + (static_cast<size_t>(k0) * kSliceRows + l) * B;
    Acc acc[KC];
#pragma unroll
    for (int w = 0; w < KC; ++w) acc[w] = Acc(0);
    // two slots in flight (C5, 8 columns: 153.4 -> 145.9 us; four: 155.5)
#pragma unroll 2
    for (int k = 0; k < km; ++k) {
      int J;
      if (rel) {
        J = node + ro[k0 + k];
        J = J < 0 ? 0 : (J > last ? last : J);  // outside the operator: zero values
      } else if (S.sidx16) {
        J = node0 + static_cast<int>(S.sidx16[static_cast<size_t>(k0 + k) * NPS + (node - node0)]);
      } else {
        J = S.sidx[static_cast<size_t>(k0 + k) * NPS + (node - node0)];
      }
      const T* __restrict__ vk = v + static_cast<size_t>(k) * kSliceRows * B;
      const T* __restrict__ xj = X + static_cast<size_t>(B * J) * ldk + c0;
#pragma unroll
      for (int c = 0; c < B; ++c) {
        T x[KC];
        vload<T, KC>(xj + static_cast<size_t>(c) * ldk, x);
        const Acc a = static_cast<Acc>(vk[c]);
#pragma unroll
        for (int w = 0; w < KC; ++w) acc[w] += a * static_cast<Acc>(x[w]);
      }
    }
    if (g < nrows) {
      const size_t idx = static_cast<size_t>(g) * ldk + c0;
      if constexpr (MODE == 0) {
        T xo[KC], yo[KC];
        vload<T, KC>(X + idx, xo);
#pragma unroll
        for (int w = 0; w < KC; ++w) yo[w] = static_cast<T>(acc[w]);
        if (dblk) {  // FP64 corrected diagonal block: y = acc + D_II x_I in FP64
          const int r = g - node * B;
          double dsum[KC];
#pragma unroll
          for (int w = 0; w < KC; ++w) dsum[w] = static_cast<double>(acc[w]);
#pragma unroll
          for (int c = 0; c < B; ++c) {
            T xi[KC];
            vload<T, KC>(X + static_cast<size_t>(B * node + c) * ldk + c0, xi);
            const double dv = dblk[(static_cast<size_t>(node) * B + r) * B + c];
#pragma unroll
            for (int w = 0; w < KC; ++w) dsum[w] += dv * static_cast<double>(xi[w]);
          }
#pragma unroll
          for (int w = 0; w < KC; ++w) yo[w] = static_cast<T>(dsum[w]);
        }
        vstore<T, KC>(Y + idx, yo);
#pragma unroll
        for (int w = 0; w < KC; ++w)
          dot[w] += static_cast<double>(xo[w]) * static_cast<double>(yo[w]);
      } else {
        double f[KC];
        T y[KC];
        vload<double, KC>(F + idx, f);
#pragma unroll
        for (int w = 0; w < KC; ++w) {
          const double yy = f[w] - static_cast<double>(acc[w]);
          y[w] = static_cast<T>(yy);
          dot[w] += yy * yy;
        }
        vstore<T, KC>(Y + idx, y);
      }
    }
  }
  // per-column sums over the workgroup (every thread holds all KC columns)
#pragma unroll
  for (int w = 0; w < KC; ++w) lds[threadIdx.x * KC + w] = dot[w];
  __syncthreads();
  if (static_cast<int>(threadIdx.x) < KC) {
    const int t = threadIdx.x;
    double sum = 0.0;
    for (int th = 0; th < kBT; ++th) sum += lds[th * KC + t];
    part[static_cast<size_t>(blockIdx.y * KC + t) * kMaxGrid + blockIdx.x] = sum;
  }
}

// ---- SpMM dispatch ----------------------------------------------------------------------------
template <typename T, int B, int MODE>
void spmm_b(const BlkTune& t, const BlkOp& A, int ldk, const T* X, T* Y, const double* F,
            const int* act, double* part, hipStream_t s) {
  const dim3 grid(t.nrb, ldk / (t.cw * t.lg));
  if (t.stage == 3) {  // sliced ELL, KC = cw columns per lane (lg = 1)
    if constexpr (B > 1) {
      if (A.sl.nslices > 0 && A.sl.nlong == 0 && t.lg == 1) {
        if (t.cw == 8) {
          hipLaunchKernelGGL((k_blk_spmm_sell<T, B, 8, MODE>), grid, dim3(kBT), 0, s, A, ldk, t.xcd,
                             X, Y, F, act, part);
          return;
        }
        if (t.cw == 4) {
          hipLaunchKernelGGL((k_blk_spmm_sell<T, B, 4, MODE>), grid, dim3(kBT), 0, s, A, ldk, t.xcd,
                             X, Y, F, act, part);
          return;
        }
        if (t.cw == 2) {
          hipLaunchKernelGGL((k_blk_spmm_sell<T, B, 2, MODE>), grid, dim3(kBT), 0, s, A, ldk, t.xcd,
                             X, Y, F, act, part);
          return;
        }
      }
    }
    throw_error(VBH_ERR_HIP, "block SpMM: no sliced-ELL kernel for the tuned configuration");
  }
#define VBH_SPMM_CASE(CW_, LG_)                                                                 \
  if (t.cw == CW_ && t.lg == LG_) {                                                             \
    if (t.stage == 2) {                                                                         \
      if constexpr (B > 1 && has_group_bcast(LG_)) {                                            \
        hipLaunchKernelGGL((k_blk_spmm_bcast<T, B, CW_, LG_, MODE>), grid, dim3(kBT), 0, s, A,  \
                           ldk, t.xcd, X, Y, F, act, part);                                     \
        return;                                                                                 \
      }                                                                                         \
      throw_error(VBH_ERR_HIP, "block SpMM: no broadcast kernel for the tuned configuration");  \
    }                                                                                           \
    if (t.stage == 1)                                                                           \
      hipLaunchKernelGGL((k_blk_spmm_lds<T, B, CW_, LG_, MODE>), grid, dim3(kBT), 0, s, A, ldk, \
                         t.xcd, X, Y, F, act, part);                                            \
    else                                                                                        \
      hipLaunchKernelGGL((k_blk_spmm<T, B, CW_, LG_, MODE>), grid, dim3(kBT), 0, s, A, ldk,    \
                         t.xcd, X, Y, F, act, part);                                            \
    return;                                                                                     \
  }
  VBH_SPMM_CASE(1, 8)
  VBH_SPMM_CASE(1, 16)
  VBH_SPMM_CASE(1, 32)
  VBH_SPMM_CASE(1, 64)
  if constexpr (B > 1) {
    VBH_SPMM_CASE(2, 4)
    VBH_SPMM_CASE(2, 8)
    VBH_SPMM_CASE(2, 16)
    VBH_SPMM_CASE(2, 32)
    VBH_SPMM_CASE(4, 2)
    VBH_SPMM_CASE(4, 4)
    VBH_SPMM_CASE(4, 8)
    VBH_SPMM_CASE(4, 16)
    VBH_SPMM_CASE(8, 1)
    VBH_SPMM_CASE(8, 2)
    VBH_SPMM_CASE(8, 4)
    VBH_SPMM_CASE(8, 8)
  }
#undef VBH_SPMM_CASE
  throw_error(VBH_ERR_HIP, "block SpMM: no kernel for the tuned configuration");
}

template <typename T, int MODE>
void spmm(const BlkTune& t, const BlkOp& A, int ldk, const T* X, T* Y, const double* F,
          const int* act, double* part, hipStream_t s) {
  if (A.B == 6) spmm_b<T, 6, MODE>(t, A, ldk, X, Y, F, act, part, s);
  else if (A.B == 3) spmm_b<T, 3, MODE>(t, A, ldk, X, Y, F, act, part, s);
  else spmm_b<T, 1, MODE>(t, A, ldk, X, Y, F, act, part, s);
}

int pow2_at_least(int k) {
  int p = 8;  // narrowest column tile with a kernel
  while (p < k) p <<= 1;
  return p;
}

int log2_of(int p) {
  int l = 0;
  while ((1 << l) < p) ++l;
  return l;
}

// SpMM candidates for ldk columns and block size B: every (CW, LG) with a
// kernel whose tile is the full width (or half of it with CW = 4), three grids (512
// workgroups keep fewer block rows in flight per XCD, so the X rows their
// neighbours gather stay in that XCD's L2 for wide FP64 blocks).
std::vector<BlkTune> candidates(int ldk, int B, bool sell) {
  const int kt = std::min(ldk, 64);
  std::vector<BlkTune> c;
  auto add = [&](int cw, int lg) {
    for (int stage : {2, 1, 0}) {
      if (stage == 2 && (!has_group_bcast(lg) || B == 1)) continue;
      for (int xcd : {1, 0})
        for (int nrb : {2048, 1024, 512}) {
          BlkTune t;
          t.cw = cw;
          t.lg = lg;
          t.nrb = nrb;
          t.xcd = xcd;
          t.stage = stage;
          c.push_back(t);
        }
    }
  };
  for (int cw : {2, 4, 8, 1}) {
    if (B == 1 && cw > 1) continue;
    const int lg = kt / cw;
    if (lg >= 1 && lg <= 64 && (cw == 1 ? lg >= 8 : true)) add(cw, lg);
  }
  if (sell && B > 1)  // sliced ELL: 8, 4 or 2 columns per lane, one lane per row
    for (int kc : {8, 4, 2})
      for (int xcd : {1, 0})
        for (int nrb : {2048, 1024}) {
          BlkTune t;
          t.cw = kc;
          t.lg = 1;
          t.nrb = nrb;
          t.xcd = xcd;
          t.stage = 3;
          c.push_back(t);
        }
  return c;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------
static BlockState& block_state(Ctx& c) {
  if (!c.blk) c.blk.reset(new BlockState());
  return *c.blk;
}

// Array-of-blocks copy of the operator (once per operator version).
static void ensure_block_operator(Ctx& c, BlockState& b) {
  if (b.op_version == c.operator_version) return;
  hipStream_t s = c.stream;
  b.drop_graph();
  b.tuned.clear();
  if (c.format == 3 || c.format == 6) {
    const int B = c.format;
    const size_t nv = static_cast<size_t>(c.bsrd.nnzb) * B * B;
    b.aos64.resize(nv);
    b.aos32.resize(nv);
    const int grid = grid_for(c.bsrd.nbr, kBT);
    if (B == 3)
      hipLaunchKernelGGL(k_bsr_to_aos<3>, dim3(grid), dim3(kBT), 0, s, c.bsrd.nbr, c.bsrd.brp,
                         c.bsrd.v64, b.aos64.p, b.aos32.p);
    else
      hipLaunchKernelGGL(k_bsr_to_aos<6>, dim3(grid), dim3(kBT), 0, s, c.bsrd.nbr, c.bsrd.brp,
                         c.bsrd.v64, b.aos64.p, b.aos32.p);
    b.op.B = B;
    b.op.nbr = c.bsrd.nbr;
    b.op.brp = c.bsrd.brp;
    b.op.bci = c.bsrd.bci;
    b.op.v64 = b.aos64.p;
    b.op.v32 = b.aos32.p;
    b.op.dblk = c.bsrd.dblk;
    b.op.sl = c.bsrd;
  } else {
    // scalar CSR: blocks of 1 x 1, values already in array-of-blocks order
    b.aos64.release();
    b.aos32.release();
    b.op.B = 1;
    b.op.nbr = c.csr.n;
    b.op.brp = c.csr.rp;
    b.op.bci = c.csr.ci;
    b.op.v64 = c.csr.v64;
    b.op.v32 = c.csr.v32;
    b.op.dblk = nullptr;
    b.op.sl = BsrDev{};
  }
  VBH_HIP(hipStreamSynchronize(s));
  b.op_version = c.operator_version;
}

static void ensure_block_work(Ctx& c, BlockState& b, int k, int precision) {
  if (b.wide.on) {  // an earlier solve stopped inside a compacted layout
    b.wide.reset();
    b.ldk = -1;
  }
  const int ldk = pow2_at_least(k);
  const int npad = c.n_pad;
  if (!(b.ldk == ldk && b.npad == npad && b.work_version == c.operator_version)) {
    const size_t m = static_cast<size_t>(npad) * ldk;
    b.F.resize(m);
    b.u.resize(m);
    b.r64.resize(m);
    b.rd.release();
    b.pd.release();
    b.Apd.release();
    b.rf.release();
    b.pf.release();
    b.Apf.release();
    b.duf.release();
    b.partA.resize(static_cast<size_t>(ldk) * kMaxGrid);
    b.partB.resize(static_cast<size_t>(3) * ldk * kMaxGrid);
    b.partR.resize(static_cast<size_t>(ldk) * kMaxGrid);
    b.coefA.resize(static_cast<size_t>(ldk));
    b.coefB.resize(static_cast<size_t>(ldk));
    b.actD.resize(static_cast<size_t>(ldk));
    b.actU.resize(static_cast<size_t>(ldk));
    b.reload.resize(static_cast<size_t>(ldk));
    b.ones.resize(static_cast<size_t>(ldk));
    b.cols.resize(static_cast<size_t>(ldk));
    if (b.cols_host) (void)hipHostFree(b.cols_host);
    b.cols_host = nullptr;
    VBH_HIP(hipHostMalloc(reinterpret_cast<void**>(&b.cols_host), sizeof(BlkCol) * ldk, 0));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c.stream, ldk, b.ones.p, 1);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c.stream, ldk, b.reload.p, 0);
    b.ldk = ldk;
    b.lg2 = log2_of(ldk);
    b.npad = npad;
    b.work_version = c.operator_version;
    b.nvec = grid_for(static_cast<long long>(m), kBT * 8);
    b.drop_graph();
  }
  const size_t m = static_cast<size_t>(npad) * ldk;
  if (precision == VBH_PREC_FP64 && !b.rd.p) {
    b.rd.resize(m);
    b.pd.resize(m);
    b.Apd.resize(m);
  }
  if (precision == VBH_PREC_MIXED && !b.rf.p) {
    b.rf.resize(m);
    b.pf.resize(m);
    b.Apf.resize(m);
    b.duf.resize(m);
  }
  b.k = k;
}

// Autotune: time every candidate SpMM configuration for this operator, ldk
// and precision on the current work vectors; keep the fastest.
static BlkTune tune(Ctx& c, BlockState& b, int precision) {
  // keyed by the width and by the live columns rounded up to a power of two:
  // a narrow column tile only pays where few of the width's columns run
  int kq = 1;
  while (kq < b.k) kq <<= 1;
  const auto key = std::make_pair(b.ldk * 128 + kq, precision);
  auto it = b.tuned.find(key);
  if (it != b.tuned.end()) return it->second;
  {  // the candidates run with the solve's live columns (the first k) marked
    std::vector<int> live(static_cast<size_t>(b.ldk));
    for (int j = 0; j < b.ldk; ++j) live[static_cast<size_t>(j)] = j < b.k ? 1 : 0;
    VBH_HIP(hipMemcpyAsync(b.ones.p, live.data(), sizeof(int) * live.size(), hipMemcpyHostToDevice,
                           c.stream));
    VBH_HIP(hipStreamSynchronize(c.stream));  // live leaves scope
  }
  // (the sliced ELL holds whole rows only when no row overflows kSellCap)
  std::vector<BlkTune> cand = candidates(b.ldk, b.op.B, b.op.sl.nslices > 0 && b.op.sl.nlong == 0);
  // VBH_BLOCK_TUNE=0 takes the first candidate; "cw,lg,nrb[,xcd[,stage]]" forces one.
  const char* env = std::getenv("VBH_BLOCK_TUNE");
  if (env && env[0] == '0') cand.resize(1);
  if (env && std::strchr(env, ',')) {
    BlkTune f;
    if (std::sscanf(env, "%d,%d,%d,%d,%d", &f.cw, &f.lg, &f.nrb, &f.xcd, &f.stage) >= 3) {
      bool known = false;
      for (const auto& t : cand) known |= (t.cw == f.cw && t.lg == f.lg && t.stage == f.stage);
      if (!known || f.nrb < 8 || f.nrb > kMaxGrid || f.nrb % 8)
        throw_error(VBH_ERR_ARG, "VBH_BLOCK_TUNE: no such SpMM configuration for this width");
      cand.assign(1, f);
    }
  }
  hipStream_t s = c.stream;
  // inputs: X = row-major F (finite values), output into the Ap buffer
  auto run = [&](const BlkTune& t) {
    if (precision == VBH_PREC_FP64)
      spmm<double, 0>(t, b.op, b.ldk, b.F.p, b.Apd.p, nullptr, b.ones.p, b.partA.p, s);
    else
      spmm<float, 0>(t, b.op, b.ldk, b.pf.p, b.Apf.p, nullptr, b.ones.p, b.partA.p, s);
  };
  if (precision == VBH_PREC_MIXED) {
    const long long m = static_cast<long long>(b.npad) * b.ldk;
    hipLaunchKernelGGL(k_blk_init<float>, dim3(b.nvec), dim3(kBT), 0, s, m, b.ldk, b.lg2, b.F.p,
                       c.diag32.p, b.u.p, b.pf.p, b.Apf.p, b.partB.p);
  }
  const char* log_env = std::getenv("VBH_BLOCK_TUNE_LOG");
  const bool log = log_env && log_env[0] == '1';
  BlkTune best;
  best.ms = 1e300;
  for (BlkTune t : cand) {
    run(t);  // warm-up (instruction cache, L2)
    VBH_HIP(hipEventRecord(c.ev_start, s));
    const int reps = 3;
    for (int i = 0; i < reps; ++i) run(t);
    VBH_HIP(hipEventRecord(c.ev_stop, s));
    VBH_HIP(hipEventSynchronize(c.ev_stop));
    float ms = 0.0f;
    VBH_HIP(hipEventElapsedTime(&ms, c.ev_start, c.ev_stop));
    t.ms = static_cast<double>(ms) / reps;
    if (log)
      std::fprintf(stderr,
                   "vbh block tune: B=%d ldk=%d prec=%d cw=%d lg=%d nrb=%d xcd=%d stage=%d  %.4f ms\n",
                   b.op.B, b.ldk, precision, t.cw, t.lg, t.nrb, t.xcd, t.stage, t.ms);
    if (t.ms < best.ms) best = t;
  }
  b.tuned[key] = best;
  return best;
}

// settle = 1: the last update of each column (still pending in actU) and
// the FP32 increments into u, then actU cleared; settle = 0: only du (mixed)
// and actU cleared, at the start of a solve or of a compacted layout.
static void settle_batch(Ctx& c, BlockState& b, int precision, int settle) {
  hipStream_t s = c.stream;
  const long long m = static_cast<long long>(b.npad) * b.ldk;
  const dim3 grid(grid_for(m, kBT * 4));
  if (precision == VBH_PREC_MIXED)
    hipLaunchKernelGGL(k_blk_settle<float>, grid, dim3(kBT), 0, s, m, b.ldk, settle, b.u.p,
                       b.duf.p, b.pf.p, b.coefA.p, b.actU.p);
  else if (settle)
    hipLaunchKernelGGL(k_blk_settle<double>, grid, dim3(kBT), 0, s, m, b.ldk, settle, b.u.p,
                       static_cast<float*>(nullptr), b.pd.p, b.coefA.p, b.actU.p);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, b.ldk, b.actU.p, 0);
}

static void enqueue_block_iterations(Ctx& c, BlockState& b, int precision, int count) {
  hipStream_t s = c.stream;
  const long long m = static_cast<long long>(b.npad) * b.ldk;
  const PcgParams prm = c.params();
  const int mixed = precision == VBH_PREC_MIXED;
  for (int i = 0; i < count; ++i) {
    // the direction kernel also adds the previous iteration's alpha p to the
    // solution (u in FP64, the FP32 increments du in mixed precision)
    if (!mixed) {
      hipLaunchKernelGGL((k_blk_direction<double, double>), dim3(b.nvec), dim3(kBT), 0, s, m,
                         b.ldk, b.lg2, c.diag64.p, b.rd.p, b.pd.p, b.coefB.p, b.actD.p, b.u.p,
                         b.coefA.p, b.actU.p);
      spmm<double, 0>(b.cur, b.op, b.ldk, b.pd.p, b.Apd.p, nullptr, b.actD.p, b.partA.p, s);
    } else {
      hipLaunchKernelGGL((k_blk_direction<float, float>), dim3(b.nvec), dim3(kBT), 0, s, m, b.ldk,
                         b.lg2, c.diag32.p, b.rf.p, b.pf.p, b.coefB.p, b.actD.p, b.duf.p,
                         b.coefA.p, b.actU.p);
      spmm<float, 0>(b.cur, b.op, b.ldk, b.pf.p, b.Apf.p, nullptr, b.actD.p, b.partA.p, s);
    }
    hipLaunchKernelGGL(k_blk_scal_alpha, dim3(b.ldk), dim3(kBT), 0, s, b.ldk, b.partA.p,
                       b.cur.nrb, b.cols.p, b.coefA.p, b.actU.p, b.actD.p);
    if (!mixed)
      hipLaunchKernelGGL(k_blk_update<double>, dim3(b.nvec), dim3(kBT), 0, s, m, b.ldk, b.lg2,
                         c.diag64.p, b.rd.p, b.Apd.p, b.coefA.p, b.actU.p, b.partB.p);
    else
      hipLaunchKernelGGL(k_blk_update<float>, dim3(b.nvec), dim3(kBT), 0, s, m, b.ldk, b.lg2,
                         c.diag32.p, b.rf.p, b.Apf.p, b.coefA.p, b.actU.p, b.partB.p);
    hipLaunchKernelGGL(k_blk_scal_beta, dim3(b.ldk), dim3(kBT), 0, s, 1, b.ldk, b.k, mixed,
                       b.partB.p, b.nvec, b.cols.p, b.coefB.p, b.actD.p, b.reload.p, prm);
  }
  settle_batch(c, b, precision, 1);
}

static hipGraphExec_t block_graph(Ctx& c, BlockState& b, int precision) {
  if (b.graph && b.g_prec == precision && b.g_ipg == c.iters_per_graph && b.g_tol == c.tol &&
      b.g_max_iters == c.max_iters && b.g_delta == c.delta() && b.g_cw == b.cur.cw &&
      b.g_lg == b.cur.lg && b.g_nrb == b.cur.nrb && b.g_xcd == b.cur.xcd &&
      b.g_stage == b.cur.stage)
    return b.graph;
  b.drop_graph();
  hipGraph_t graph = nullptr;
  VBH_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  enqueue_block_iterations(c, b, precision, c.iters_per_graph);
  VBH_HIP(hipStreamEndCapture(c.stream, &graph));
  VBH_HIP(hipGraphInstantiate(&b.graph, graph, nullptr, nullptr, 0));
  (void)hipGraphDestroy(graph);
  b.g_prec = precision;
  b.g_ipg = c.iters_per_graph;
  b.g_tol = c.tol;
  b.g_max_iters = c.max_iters;
  b.g_delta = c.delta();
  b.g_cw = b.cur.cw;
  b.g_lg = b.cur.lg;
  b.g_nrb = b.cur.nrb;
  b.g_xcd = b.cur.xcd;
  b.g_stage = b.cur.stage;
  return b.graph;
}

static void read_cols(Ctx& c, BlockState& b) {
  VBH_HIP(hipMemcpyAsync(b.cols_host, b.cols.p, sizeof(BlkCol) * b.ldk, hipMemcpyDeviceToHost,
                         c.stream));
  VBH_HIP(hipStreamSynchronize(c.stream));
}

// ---- column compaction -------------------------------------------------------------------
static void swap_work(BlockState& b, BlockState::Wide& w) {
  b.F.swap(w.F);
  b.u.swap(w.u);
  b.r64.swap(w.r64);
  b.rd.swap(w.rd);
  b.pd.swap(w.pd);
  b.Apd.swap(w.Apd);
  b.partA.swap(w.partA);
  b.partB.swap(w.partB);
  b.partR.swap(w.partR);
  b.coefA.swap(w.coefA);
  b.coefB.swap(w.coefB);
  b.rf.swap(w.rf);
  b.pf.swap(w.pf);
  b.Apf.swap(w.Apf);
  b.duf.swap(w.duf);
  b.actD.swap(w.actD);
  b.actU.swap(w.actU);
  b.reload.swap(w.reload);
  b.ones.swap(w.ones);
  b.cols.swap(w.cols);
  std::swap(b.ldk, w.ldk);
  std::swap(b.lg2, w.lg2);
  std::swap(b.k, w.k);
  std::swap(b.nvec, w.nvec);
}

// Finished slots of the current layout: final state into res, solution into
// the full-width buffer (the layout in use after the first compaction).
static void retire_columns(Ctx& c, BlockState& b, const std::vector<int>& orig,
                           std::vector<BlkCol>& res, bool all) {
  std::vector<int> ret(static_cast<size_t>(b.ldk), -1);
  for (int j = 0; j < b.ldk; ++j)
    if ((all || b.cols_host[j].state != kColActive) && orig[static_cast<size_t>(j)] >= 0) {
      res[static_cast<size_t>(orig[static_cast<size_t>(j)])] = b.cols_host[j];
      ret[static_cast<size_t>(j)] = orig[static_cast<size_t>(j)];
    }
  if (!b.wide.on) return;  // the full-width layout is the current one
  DBuf<int> dret;
  dret.upload(ret.data(), ret.size(), c.stream);
  const long long m = static_cast<long long>(b.npad) * b.ldk;
  hipLaunchKernelGGL(k_blk_scatter<double>, dim3(grid_for(m, kBT * 4)), dim3(kBT), 0, c.stream,
                     static_cast<long long>(b.npad), b.lg2, b.wide.ldk, dret.p, b.u.p, b.wide.u.p);
  VBH_HIP(hipStreamSynchronize(c.stream));
}

// Moves the running columns (all active, none paused) into a layout of
// pow2_at_least(count) columns. orig[j] is the caller's column of slot j.
static void compact_columns(Ctx& c, BlockState& b, int precision, std::vector<int>& orig,
                            std::vector<BlkCol>& res) {
  hipStream_t s = c.stream;
  std::vector<int> live;
  for (int j = 0; j < b.ldk; ++j)
    if (b.cols_host[j].state == kColActive) live.push_back(j);
  const int k2 = static_cast<int>(live.size());
  const int ldk2 = pow2_at_least(k2);
  retire_columns(c, b, orig, res, false);
  std::vector<double> cb(static_cast<size_t>(b.ldk));
  VBH_HIP(hipMemcpyAsync(cb.data(), b.coefB.p, sizeof(double) * b.ldk, hipMemcpyDeviceToHost, s));
  VBH_HIP(hipStreamSynchronize(s));
  // the source layout: the full-width buffers on the first compaction (kept
  // aside for the results), otherwise the current narrow ones (freed after)
  BlockState::Wide cur;
  const bool first = !b.wide.on;
  if (first) {
    swap_work(b, b.wide);
    b.wide.on = true;
  } else {
    swap_work(b, cur);
  }
  const BlockState::Wide& src = first ? b.wide : cur;
  const size_t m2 = static_cast<size_t>(b.npad) * ldk2;
  {
    BlockState::Wide nb;
    nb.ldk = ldk2;
    nb.lg2 = log2_of(ldk2);
    nb.k = k2;
    nb.nvec = grid_for(static_cast<long long>(m2), kBT * 8);
    nb.F.resize(m2);
    nb.u.resize(m2);
    nb.r64.resize(m2);
    if (precision == VBH_PREC_MIXED) {
      nb.rf.resize(m2);
      nb.pf.resize(m2);
      nb.Apf.resize(m2);
      nb.duf.resize(m2);
    } else {
      nb.rd.resize(m2);
      nb.pd.resize(m2);
      nb.Apd.resize(m2);
    }
    nb.partA.resize(static_cast<size_t>(ldk2) * kMaxGrid);
    nb.partB.resize(static_cast<size_t>(3) * ldk2 * kMaxGrid);
    nb.partR.resize(static_cast<size_t>(ldk2) * kMaxGrid);
    swap_work(b, nb);  // b.* was emptied above
  }
  std::vector<int> g(static_cast<size_t>(ldk2), -1);
  for (int j = 0; j < k2; ++j) g[static_cast<size_t>(j)] = live[static_cast<size_t>(j)];
  DBuf<int> dg;
  dg.upload(g.data(), g.size(), s);
  const dim3 grid(grid_for(static_cast<long long>(m2), kBT * 4));
  auto move64 = [&](DBuf<double>& dst, const DBuf<double>& from) {
    hipLaunchKernelGGL(k_blk_gather<double>, grid, dim3(kBT), 0, s, static_cast<long long>(b.npad),
                       src.ldk, b.lg2, dg.p, from.p, dst.p);
  };
  auto move32 = [&](DBuf<float>& dst, const DBuf<float>& from) {
    hipLaunchKernelGGL(k_blk_gather<float>, grid, dim3(kBT), 0, s, static_cast<long long>(b.npad),
                       src.ldk, b.lg2, dg.p, from.p, dst.p);
  };
  // per-column state: the running columns' as they are, padding columns done
  std::vector<double> cb2(static_cast<size_t>(ldk2), 0.0), zero64(static_cast<size_t>(ldk2), 0.0);
  std::vector<int> on(static_cast<size_t>(ldk2), 0), zero(static_cast<size_t>(ldk2), 0),
      one(static_cast<size_t>(ldk2), 1);
  std::vector<BlkCol> cols2(static_cast<size_t>(ldk2));
  std::vector<int> orig2(static_cast<size_t>(ldk2), -1);
  for (int j = 0; j < ldk2; ++j) {
    BlkCol& cj = cols2[static_cast<size_t>(j)];
    if (j < k2) {
      const int q = live[static_cast<size_t>(j)];
      cj = b.cols_host[q];
      cb2[static_cast<size_t>(j)] = cb[static_cast<size_t>(q)];
      on[static_cast<size_t>(j)] = 1;
      orig2[static_cast<size_t>(j)] = orig[static_cast<size_t>(q)];
    } else {
      std::memset(&cj, 0, sizeof cj);
      cj.state = kColDone;
      cj.flags = kFlagZeroRhs;
      cj.norm_b = 1.0;
    }
  }
  b.ones.upload(one.data(), one.size(), s);
  move64(b.F, src.F);
  // the SpMM shape for this width: tuning (first time only) runs its
  // candidates on these buffers and overwrites u, p, Ap and the partials, so
  // the running columns' vectors are moved in after it
  b.drop_graph();
  b.cur = tune(c, b, precision);
  move64(b.u, src.u);
  if (precision == VBH_PREC_MIXED) {
    move32(b.rf, src.rf);
    move32(b.pf, src.pf);
  } else {
    move64(b.rd, src.rd);
    move64(b.pd, src.pd);
  }
  b.coefA.upload(zero64.data(), zero64.size(), s);
  b.coefB.upload(cb2.data(), cb2.size(), s);
  b.actD.upload(on.data(), on.size(), s);
  b.actU.upload(zero.data(), zero.size(), s);
  // the last batch settled every pending update into u: the new du starts clear
  settle_batch(c, b, precision, 0);
  b.reload.upload(zero.data(), zero.size(), s);
  b.cols.upload(cols2.data(), cols2.size(), s);
  VBH_HIP(hipStreamSynchronize(s));
  orig = orig2;
}

// After the solve: every column's state and solution back in the full-width
// layout the caller reads.
static void finish_compaction(Ctx& c, BlockState& b, int precision, std::vector<int>& orig,
                              std::vector<BlkCol>& res) {
  if (!b.wide.on) return;
  read_cols(c, b);
  retire_columns(c, b, orig, res, true);
  BlockState::Wide cur;
  swap_work(b, cur);
  swap_work(b, b.wide);
  b.wide.on = false;
  b.cols.upload(res.data(), res.size(), c.stream);
  VBH_HIP(hipStreamSynchronize(c.stream));
  b.drop_graph();
  b.cur = tune(c, b, precision);  // the full-width shape (cached), as reported
}

// Runs the block solve on b.F (row-major) into b.u.
static void run_block_solve(Ctx& c, BlockState& b, int precision) {
  hipStream_t s = c.stream;
  long long m = static_cast<long long>(b.npad) * b.ldk;
  const PcgParams prm = c.params();
  const int mixed = precision == VBH_PREC_MIXED;
  b.cur = tune(c, b, precision);
  // VBH_BLOCK_COMPACT=0 keeps the full width; so does a forced SpMM shape
  // (VBH_BLOCK_TUNE=cw,lg,...), which exists for this width only
  const char* cp_env = std::getenv("VBH_BLOCK_COMPACT");
  const char* tn_env = std::getenv("VBH_BLOCK_TUNE");
  const bool compact = !(cp_env && cp_env[0] == '0') && !(tn_env && std::strchr(tn_env, ','));
  std::vector<int> orig(static_cast<size_t>(b.ldk));
  for (int j = 0; j < b.ldk; ++j) orig[static_cast<size_t>(j)] = j;
  std::vector<BlkCol> res(static_cast<size_t>(b.ldk));
  const char* lg_env = std::getenv("VBH_BLOCK_LOG");
  const bool blog = lg_env && lg_env[0] == '1';
  long long n_refresh_passes = 0, n_launches = 0;
  DBuf<int> pmask;
  VBH_HIP(hipEventRecord(c.ev_start, s));
  if (!mixed)
    hipLaunchKernelGGL(k_blk_init<double>, dim3(b.nvec), dim3(kBT), 0, s, m, b.ldk, b.lg2, b.F.p,
                       c.diag64.p, b.u.p, b.rd.p, b.pd.p, b.partB.p);
  else
    hipLaunchKernelGGL(k_blk_init<float>, dim3(b.nvec), dim3(kBT), 0, s, m, b.ldk, b.lg2, b.F.p,
                       c.diag32.p, b.u.p, b.rf.p, b.pf.p, b.partB.p);
  settle_batch(c, b, precision, 0);  // no update pending, du clear
  hipLaunchKernelGGL(k_blk_scal_beta, dim3(b.ldk), dim3(kBT), 0, s, 0, b.ldk, b.k, mixed,
                     b.partB.p, b.nvec, b.cols.p, b.coefB.p, b.actD.p, b.reload.p, prm);
  const char* ge_env = std::getenv("VBH_USE_GRAPH");
  const bool use_graph = !(ge_env && ge_env[0] == '0');
  hipGraphExec_t ge = use_graph ? block_graph(c, b, precision) : nullptr;
  // every pass either advances an active column or refreshes a paused one
  const long long cap = 4ll * (static_cast<long long>(c.max_iters) + 2) + 64;
  long long passes = 0;
  for (;;) {
    read_cols(c, b);
    int active = 0, paused = 0;
    for (int j = 0; j < b.ldk; ++j) {
      active += b.cols_host[j].state == kColActive;
      paused += b.cols_host[j].state == kColPaused;
    }
    if (!active && !paused) break;
    if (++passes > cap) throw_error(VBH_ERR_HIP, "HIP block PCG: device state stopped advancing");
    m = static_cast<long long>(b.npad) * b.ldk;
    if (!paused && compact && b.ldk >= 16 && active <= b.ldk / 2) {
      if (blog) std::fprintf(stderr, "vbh block: launch %lld: %d running columns -> width %d\n",
                             n_launches, active, pow2_at_least(active));
      compact_columns(c, b, precision, orig, res);
      ge = use_graph ? block_graph(c, b, precision) : nullptr;
      continue;
    }
    if (paused) {
      ++n_refresh_passes;
      // r = F - K u with the tuned shape of the current width, column tiles
      // without a paused column skipped (pmask). Every pass streams the whole
      // FP64 operator, so narrower shapes for a few paused columns are not
      // cheaper (C3, one 8-column tile: 1.11 ms with direct loads in groups of
      // 8, 0.84 ms with the 4-lane broadcast; all 32 columns with the tuned
      // shape: 0.64 ms). Only paused columns are reloaded.
      std::vector<int> pm(static_cast<size_t>(b.ldk), 0);
      for (int j = 0; j < b.ldk; ++j) pm[static_cast<size_t>(j)] = b.cols_host[j].state == kColPaused;
      const BlkTune rt = b.cur;
      pmask.upload(pm.data(), pm.size(), s);
      spmm<double, 1>(rt, b.op, b.ldk, b.u.p, b.r64.p, b.F.p, pmask.p, b.partR.p, s);
      hipLaunchKernelGGL(k_blk_scal_refresh, dim3(b.ldk), dim3(kBT), 0, s, b.ldk, b.partR.p,
                         rt.nrb, b.cols.p, b.reload.p, prm);
      hipLaunchKernelGGL(k_blk_reload, dim3(b.nvec), dim3(kBT), 0, s, m, b.ldk, b.lg2, c.diag32.p,
                         b.r64.p, b.rf.p, b.reload.p, b.partB.p);
      hipLaunchKernelGGL(k_blk_scal_beta, dim3(b.ldk), dim3(kBT), 0, s, 2, b.ldk, b.k, mixed,
                         b.partB.p, b.nvec, b.cols.p, b.coefB.p, b.actD.p, b.reload.p, prm);
      continue;
    }
    ++n_launches;
    if (use_graph) VBH_HIP(hipGraphLaunch(ge, s));
    else enqueue_block_iterations(c, b, precision, c.iters_per_graph);
  }
  finish_compaction(c, b, precision, orig, res);
  if (blog)
    std::fprintf(stderr, "vbh block: %lld iteration launches, %lld refresh passes\n", n_launches,
                 n_refresh_passes);
  VBH_HIP(hipEventRecord(c.ev_stop, s));
  VBH_HIP(hipEventSynchronize(c.ev_stop));
  float ms = 0.0f;
  VBH_HIP(hipEventElapsedTime(&ms, c.ev_start, c.ev_stop));
  c.stats.solve_ms = ms;
}

static double column_rel(const BlkCol& cj, int precision) {
  if (cj.flags & kFlagZeroRhs) return 0.0;
  if (precision == VBH_PREC_MIXED && (cj.flags & (kFlagConverged | kFlagStall))) return cj.rel_true;
  return cj.rel;
}

static void block_solve(Ctx& c, int k, const double* F, double* U, bool device_ptrs, int* iters,
                        double* rel_res) {
  if (!c.have_matrix) throw_error(VBH_ERR_STATE, "vbh_solve_block: no matrix set");
  if (k <= 0 || k > 64) throw_error(VBH_ERR_ARG, "vbh_solve_block: k must be in 1..64");
  if (!F || !U) throw_error(VBH_ERR_ARG, "vbh_solve_block: null array");
  c.select_stream_device();
  BlockState& b = block_state(c);
  const int precision = c.precision;
  ensure_block_operator(c, b);
  ensure_block_work(c, b, k, precision);
  hipStream_t s = c.stream;
  // column-major API block (n_api rows) <-> operator rows (n, mapped under MPCs)
  DBuf<double> stage;
  stage.resize(static_cast<size_t>(c.n) * k);
  c.to_operator(F, device_ptrs, stage.p, k);
  const long long m = static_cast<long long>(b.npad) * b.ldk;
  hipLaunchKernelGGL(k_to_rows, dim3(grid_for(m, kBT * 4)), dim3(kBT), 0, s, c.n, k, b.npad, b.ldk,
                     stage.p, b.F.p);
  run_block_solve(c, b, precision);
  hipLaunchKernelGGL(k_to_cols, dim3(grid_for(static_cast<long long>(c.n) * k, kBT * 4)), dim3(kBT),
                     0, s, c.n, k, b.ldk, b.u.p, stage.p);
  c.from_operator(stage.p, U, device_ptrs, k);
  read_cols(c, b);
  int max_it = 0, refr = 0;
  double max_rel = 0.0;
  std::string failure;
  for (int j = 0; j < k; ++j) {
    const BlkCol& cj = b.cols_host[j];
    const double rel = column_rel(cj, precision);
    if (iters) iters[j] = cj.iters;
    if (rel_res) rel_res[j] = rel;
    max_it = std::max(max_it, cj.iters);
    refr += cj.refreshes;
    max_rel = std::max(max_rel, rel);
    if (failure.empty()) {
      if (cj.flags & kFlagRz0)
        failure = "HIP PCG: load case " + std::to_string(j) + ": non-positive initial r*z=" +
                  fmt_g(cj.rz) + " -- preconditioner is not SPD or matrix is singular";
      else if (cj.flags & (kFlagBreakdown | kFlagMaxIter))
        failure = "HIP PCG: load case " + std::to_string(j) + ": did not converge after " +
                  std::to_string(c.max_iters) + " iterations (relative residual " +
                  fmt_g((cj.flags & kFlagBreakdown) ? 1e300 : cj.rel) + " > tolerance " +
                  fmt_g(c.tol) +
                  "). Consider increasing max iterations or using a direct solver.";
    }
  }
  c.stats.iterations = max_it;
  c.stats.refinements = refr;
  c.stats.rel_res = max_rel;
  c.stats.format = c.format;
  c.last_iters = max_it;
  c.last_rel = max_rel;
  if (!failure.empty()) throw_error(VBH_ERR_SOLVER, failure);
}

}  // namespace vbh

extern "C" {

int vbh_solve_block(vbh_ctx* h, int k, const double* F, double* U, int* iters, double* rel_res) {
  if (!h) {
    vbh::set_error("vbh_solve_block: null ctx");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] { vbh::block_solve(*c, k, F, U, false, iters, rel_res); });
}

int vbh_solve_block_device(vbh_ctx* h, int k, const double* dF, double* dU, int* iters,
                           double* rel_res) {
  if (!h) {
    vbh::set_error("vbh_solve_block_device: null ctx");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] { vbh::block_solve(*c, k, dF, dU, true, iters, rel_res); });
}

int vbh_last_block_config(const vbh_ctx* h, vbh_block_config* out) {
  if (!h || !out) {
    vbh::set_error("vbh_last_block_config: null argument");
    return VBH_ERR_ARG;
  }
  const auto* c = reinterpret_cast<const vbh::Ctx*>(h);
  if (!c->blk) {
    vbh::set_error("vbh_last_block_config: no block solve yet");
    return VBH_ERR_STATE;
  }
  out->col_tile = c->blk->cur.cw * c->blk->cur.lg;
  out->cols_per_lane = c->blk->cur.cw;
  out->lanes_per_group = c->blk->cur.lg;
  out->row_blocks = c->blk->cur.nrb;
  out->xcd_rows = c->blk->cur.xcd;
  out->value_path = c->blk->cur.stage;
  out->ldk = c->blk->ldk;
  out->spmm_ms = c->blk->cur.ms;
  return VBH_OK;
}

}  // extern "C"
