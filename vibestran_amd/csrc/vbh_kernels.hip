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
#include <hip/hip_ext.h>

#include "vbh_internal.hpp"

#include <algorithm>
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

// ---- four consecutive rows per thread ----------------------------------------------
// The PCG vector kernels stream 3-8 vectors; giving each thread 4 consecutive
// rows with one 16/32-byte access per vector puts every load of the thread in
// flight at once (one memory round trip instead of one per grid-stride step).
// Work vectors come from hipMalloc (256-byte aligned), so row 4q is aligned.
template <typename T>
struct Quad {
  T v[4];
};
__device__ __forceinline__ Quad<float> ld4(const float* p) {
  const float4 a = *reinterpret_cast<const float4*>(p);
  return {{a.x, a.y, a.z, a.w}};
}
__device__ __forceinline__ Quad<double> ld4(const double* p) {
  const double2 a = *reinterpret_cast<const double2*>(p);
  const double2 b = *reinterpret_cast<const double2*>(p + 2);
  return {{a.x, a.y, b.x, b.y}};
}
__device__ __forceinline__ void st4(float* p, const Quad<float>& q) {
  *reinterpret_cast<float4*>(p) = make_float4(q.v[0], q.v[1], q.v[2], q.v[3]);
}
__device__ __forceinline__ void st4(double* p, const Quad<double>& q) {
  *reinterpret_cast<double2*>(p) = make_double2(q.v[0], q.v[1]);
  *reinterpret_cast<double2*>(p + 2) = make_double2(q.v[2], q.v[3]);
}

// ---- deterministic reductions ---------------------------------------------------
// A double moved between lanes by DPP (both 32-bit halves), CTRL one of the
// in-row patterns below: every source lane lies in the destination's row of
// 16, so masks and bound control never apply, and the move needs no old value
// (update_dpp with an old value makes the compiler initialise the destination).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

__device__ __forceinline__ double readlane_f64(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

// Sum over the 64 lanes of a wave, the same bits in every lane: a butterfly
// inside each row of 16 lanes (DPP quad_perm lane^1 and lane^2, then
// row_half_mirror and row_mirror, which swap quads and half rows), where the
// two lanes of every exchange add the same two values (x + y == y + x), then
// the four row sums (r0 + r1) + (r2 + r3). No LDS traffic, no barrier.
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_f64<0xB1>(x);   // quad_perm [1,0,3,2]
  x += dpp_f64<0x4E>(x);   // quad_perm [2,3,0,1]
  x += dpp_f64<0x141>(x);  // row_half_mirror
  x += dpp_f64<0x140>(x);  // row_mirror
  return (readlane_f64(x, 0) + readlane_f64(x, 16)) + (readlane_f64(x, 32) + readlane_f64(x, 48));
}

// Sum of v over the workgroup, returned in every thread with the same bits
// and in the same order in every workgroup (results are deterministic for a
// given grid): wave sums, then the waves' sums in wave order through LDS.
// Called by every thread of the workgroup with all lanes active; `lds` holds
// NV x (waves per block) doubles and may be reused after the call.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) lds[w * NV + i] = v[i];
  __syncthreads();
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double t = lds[i];
    for (int k = 1; k < nw; ++k) t += lds[k * NV + i];
    v[i] = t;
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

template <typename T>
__device__ __forceinline__ const T* svals(const BsrDev& A);
template <>
__device__ __forceinline__ const double* svals<double>(const BsrDev& A) { return A.s64; }
template <>
__device__ __forceinline__ const float* svals<float>(const BsrDev& A) { return A.s32; }

// Slots 0..km-1 of one lane: values at v + k * 64 * B, block column j0 + ix[k * NPS].
template <typename T, int B, typename Acc, typename IX>
__device__ __forceinline__ Acc sell_row_ix(const T* __restrict__ v, const T* __restrict__ x,
                                           const IX* __restrict__ ix, int j0, int km) {
  constexpr int NPS = sell_nps(B);
  Acc acc = Acc(0);
#pragma unroll 4
  for (int k = 0; k < km; ++k) {
    const int J = j0 + static_cast<int>(ix[k * NPS]);
    const T* vk = v + static_cast<size_t>(k) * kSliceRows * B;
#pragma unroll
    for (int c = 0; c < B; ++c) acc += static_cast<Acc>(vk[c]) * static_cast<Acc>(x[B * J + c]);
  }
  return acc;
}

template <typename T>
__device__ __forceinline__ const T* rel_vals(const BsrDev& A);
template <>
__device__ __forceinline__ const double* rel_vals<double>(const BsrDev& A) { return A.rv64; }
template <>
__device__ __forceinline__ const float* rel_vals<float>(const BsrDev& A) { return A.rv32; }

// The block slots of slice s in the stream of value type T: [k0, k0 + km)
// (scalar loads of the slice offsets).
struct Slots {
  int k0, km;
};
template <typename T>
__device__ __forceinline__ Slots sell_slots(const BsrDev& A, int s) {
  s = __builtin_amdgcn_readfirstlane(s);
  const int* __restrict__ k = A.rk64 ? (sizeof(T) == 4 ? A.rk32 : A.rk64) : A.koff;
  const int k0 = k[s];
  return {k0, k[s + 1] - k0};
}

// Row g = 64 s + l of the relative sliced ELL (BsrDev::rk64 ...) times x: the
// slot offsets are uniform over the wave's slice (scalar loads); block
// columns outside the operator (a boundary node without that neighbour,
// whose value block is zero) are clamped into range.
template <typename T, int B, typename Acc>
__device__ __forceinline__ Acc sell_row_rel(const BsrDev& A, const T* __restrict__ x, int s,
                                            int l, Slots sl) {
  const bool f32 = sizeof(T) == 4;
  const int* __restrict__ ro = f32 ? A.ro32 : A.ro64;
  const int k0 = sl.k0, km = sl.km;
  const int node = (kSliceRows * s + l) / B;
  const int last = A.nbr - 1;
  const T* __restrict__ v = rel_vals<T>(A) + (static_cast<size_t>(k0) * kSliceRows + l) * B;
  const int* __restrict__ off = ro + k0;
  Acc acc = Acc(0);
#pragma unroll 4
  for (int k = 0; k < km; ++k) {
    int J = node + off[k];
    J = J < 0 ? 0 : (J > last ? last : J);
    const T* vk = v + static_cast<size_t>(k) * kSliceRows * B;
#pragma unroll
    for (int c = 0; c < B; ++c) acc += static_cast<Acc>(vk[c]) * static_cast<Acc>(x[B * J + c]);
  }
  return acc;
}

// Row g = 64 s + l of the sliced-ELL operator times x (one lane, no reduction),
// accumulated in Acc; sl = sell_slots<T>(A, s).
template <typename T, int B, typename Acc = T>
__device__ __forceinline__ Acc sell_row(const BsrDev& A, const T* __restrict__ val,
                                        const T* __restrict__ x, int s, int l, Slots sl) {
  if (A.rk64) return sell_row_rel<T, B, Acc>(A, x, s, l, sl);
  constexpr int NPS = sell_nps(B);
  const int k0 = sl.k0, km = sl.km;
  const int node = (kSliceRows * s + l) / B;
  const int node0 = (kSliceRows * s) / B;
  const T* v = val + (static_cast<size_t>(k0) * kSliceRows + l) * B;
  if (A.sidx16)  // 16-bit offsets from node0
    return sell_row_ix<T, B, Acc>(v, x, A.sidx16 + static_cast<size_t>(k0) * NPS + (node - node0),
                                  node0, km);
  return sell_row_ix<T, B, Acc>(v, x, A.sidx + static_cast<size_t>(k0) * NPS + (node - node0), 0,
                                km);
}

template <typename T, int B, typename Acc = T>
__device__ __forceinline__ Acc sell_row(const BsrDev& A, const T* __restrict__ val,
                                        const T* __restrict__ x, int s, int l) {
  return sell_row<T, B, Acc>(A, val, x, s, l, sell_slots<T>(A, s));
}

// FP64 diagonal node block of row g (row-sum preserving FP32 stream, see
// BsrDev::dblk) times x.
template <typename T, int B>
__device__ __forceinline__ double diag_block_row(const double* __restrict__ dblk,
                                                 const T* __restrict__ x, int g) {
  const int node = g / B, r = g - node * B;
  const double* d = dblk + (static_cast<size_t>(node) * B + r) * B;
  double acc = 0.0;
#pragma unroll
  for (int c = 0; c < B; ++c) acc += d[c] * static_cast<double>(x[B * node + c]);
  return acc;
}

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

template <typename T, int B, typename Acc = T>
__global__ __launch_bounds__(kBlock) void k_spmv_bsr(BsrDev A, const T* __restrict__ x,
                                                     T* __restrict__ y) {
  const T* val = svals<T>(A);
  const int l = threadIdx.x & (kSliceRows - 1);
  const int nrows = A.nbr * B;
  const bool split = (sizeof(T) == 4) && A.dblk != nullptr;
  for (int s = (blockIdx.x * blockDim.x + threadIdx.x) / kSliceRows; s < A.nslices;
       s += gridDim.x * (blockDim.x / kSliceRows)) {
    const Acc acc = sell_row<T, B, Acc>(A, val, x, s, l);
    const int g = kSliceRows * s + l;
    if (g < nrows)
      y[g] = split ? static_cast<T>(static_cast<double>(acc) + diag_block_row<T, B>(A.dblk, x, g))
                   : static_cast<T>(acc);
  }
}

// ---- Jacobi-PCG: two kernels per iteration ---------------------------------------------
// The reference's recurrence (cuda_pcg_backend_impl.hpp:1867-1954) with every
// global sum produced as per-block partials and reduced by the NEXT kernel
// (every block redundantly, fixed order): deterministic, no atomics, no
// in-kernel fences, scalars bit-identical across blocks. The search direction
// of the next iteration is formed by the vector kernel of this one:
//
//   init       u = 0, r = F, p = r/diag; partials (r.z, r.r, F.F)            -> partB
//   cg_spmv    Ap = K p; partials (p.Ap, z.Ap, Ap.diag^-1.Ap), z = r/diag
//              formed in registers                                           -> partA
//              its last workgroup (one slice per wave: it has time to spare)
//              reduces partB                                                 -> sumB
//   cg_vector  rz_i = r.z, rr_i = r.r of the current r from sumB; the
//              convergence / pause / limit decision (the reference's
//              finalize step, impl:704-738); reduce partA ->
//              alpha = rz_i / p.Ap (breakdown if p.Ap <= 0),
//              rz_{i+1} = rz_i - 2 alpha z.Ap + alpha^2 Ap.diag^-1.Ap (the
//              value of (r - alpha Ap).diag^-1.(r - alpha Ap)),
//              beta = rz_{i+1} / rz_i;
//              u += alpha p (FP64); r -= alpha Ap; p = r/diag + beta p;
//              partials (r.z, r.r) of the new r                              -> partB
//
// z = r/diag is never stored: both kernels form it from r (the same FP
// operation, so the same value).
//
// alpha always takes r.z summed over the vectors themselves; only beta takes
// the expanded rz_{i+1}. (A CPU model of the 605k-DOF hex block: 1,639 mixed
// iterations so, 2,465 with the expansion in alpha as well, 1,657 for the
// three-kernel form that reduced r.z before forming p.)
//
// Mixed precision refreshes the residual inside the same two-kernel sequence,
// with no host round trip. cg_vector pauses, before its update, when the
// recursive residual fell by refresh_delta since the last refresh (or below
// tol), and sets refresh_req; then
//   cg_spmv    (refresh_req set) r64 = F - K64 u over the FP64 operator;
//              partial |r64|^2 in partA slot 0; spmv_mode = 1
//   cg_vector  (spmv_mode set) decide from the FP64 true residual: converged /
//              stalled, or continue with r = (float) r64, partials (r.z, r.r);
//              refresh_req = 0, the paused iteration marked resumed
// and the next cg_spmv recomputes K p with p kept (reliable update, van der
// Vorst and Ye: only r is replaced), so its z.Ap uses the refreshed z, and
// cg_vector resumes the paused iteration.
//
// Only block 0 writes the scalar state, and never a field that other blocks
// of the same kernel read: cg_vector reads the iteration as iter_cur and the
// refresh step as spmv_mode, which cg_spmv copies from iter_done and
// refresh_req (written by cg_vector); norm_b and rr_refresh are written only
// in iterations whose decision does not read them; done is set only together
// with a stop every block of the kernel reaches on its own; sumB is written
// by the SpMV's last workgroup only. partB has two buffers: the SpMV launched
// at position j of a launch sequence (a captured graph, the directly launched
// iterations) reduces buffer j & 1 and the vector kernel at position j writes
// the other; cg_init writes buffer 0, which position 0 reduces, and every
// sequence has an even length.

template <int NV>
__device__ __forceinline__ void write_partials(double (&v)[NV], double* __restrict__ part,
                                               double* lds) {
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) part[i * kMaxGrid + blockIdx.x] = v[i];
}

constexpr int kLds = 8 * kMaxRed;

// The state words the SpMV kernels branch on (done, iter_done, refresh_req),
// read by every wave itself: uniform addresses, so scalar loads, and no
// barrier. They are written only by the vector kernels, never while an SpMV
// runs.
struct SpmvWords {
  int done, iter_done, refresh_req;
};
__device__ __forceinline__ SpmvWords spmv_words(const PcgScalars* st) {
  return {st->done, st->iter_done, st->refresh_req};
}

// The head of the vector kernel, split in two so the caller can put its own
// stream loads between them:
//   head_issue   every load the head needs, all in flight at once: this
//                thread's partial sums of both producers (unrolled, so one
//                memory round trip instead of one per grid-stride step) and
//                the state words (thread 0);
//   head_finish  one block sum of all partials and the broadcast of the words.
// Partials are issued before any stream load, so they return first (vector
// loads complete in order) and are read while they are still in L2.
// Only the state words go through the head: the double scalars (norm_b,
// rr_refresh) are read after the decision that needs them.
struct Head {
  int done, iter_cur, resume_iter, spmv_mode;
};

// NJA partial sums per thread for the SpMV's grid: the smallest NJ with
// NJ * kBlock >= the grid (1, 2, 4, 8).
template <int NJA>
struct HeadLoads {
  double a[3][NJA];  // p.Ap, z.Ap, Ap.diag^-1.Ap (or |F - K u|^2 in slot 0)
  double b[3];       // sumB: r.z, r.r, F.F
  Head st;
};

template <int NJA>
__device__ __forceinline__ void head_issue(const PcgScalars* st, const CgRed& red,
                                           HeadLoads<NJA>& hl) {
  // every slot of the kMaxGrid-wide partial arrays is allocated, so the loads
  // are unconditional (no branch between them) and the unused ones masked
#pragma unroll
  for (int j = 0; j < NJA; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) hl.a[i][j] = red.partA[i * kMaxGrid + threadIdx.x + j * kBlock];
#pragma unroll
  for (int j = 0; j < NJA; ++j)
    if (threadIdx.x + j * kBlock >= red.nA)
#pragma unroll
      for (int i = 0; i < 3; ++i) hl.a[i][j] = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) hl.b[i] = st->sumB[i];  // uniform: scalar loads
  if (threadIdx.x == 0) {
    hl.st.done = st->done;
    hl.st.iter_cur = st->iter_cur;
    hl.st.resume_iter = st->resume_iter;
    hl.st.spmv_mode = st->spmv_mode;
  }
}

// out[0..2] = partA sums, out[3..5] = sumB, the same in every thread
template <int NJA>
__device__ __forceinline__ void head_finish(HeadLoads<NJA>& hl, double (&out)[6], Head& h,
                                            double* lds, Head* sh) {
  double a[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    a[i] = 0.0;
#pragma unroll
    for (int j = 0; j < NJA; ++j) a[i] += hl.a[i][j];
  }
  if (threadIdx.x == 0) *sh = hl.st;
  block_sum<3>(a, lds);  // its barriers also publish *sh
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    out[i] = a[i];
    out[3 + i] = hl.b[i];
  }
  h = *sh;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_init(int n, CgBufs<T> b, PcgScalars* st,
                                                    CgRed red) {
  __shared__ double lds[kLds];
  double v[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double f = b.F[i];
    const T ri = static_cast<T>(f);
    const T zi = ri / b.diag[i];
    b.u64[i] = 0.0;
    b.r[i] = ri;
    b.p[i] = zi;
    v[0] += static_cast<double>(ri) * static_cast<double>(zi);
    v[1] += static_cast<double>(ri) * static_cast<double>(ri);
    v[2] += f * f;
  }
  write_partials<3>(v, red.partB, lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->iter_done = 0;
    st->iter_cur = 0;
    st->resume_iter = -1;
    st->refresh_req = 0;
    st->spmv_mode = 0;
    st->flags = 0;
    st->done = 0;
    st->refreshes = 0;
    st->hist_stalls[0] = st->hist_stalls[1] = 0;
    st->hist_rel[0] = st->hist_rel[1] = 1e300;
    st->rz[0] = st->rz[1] = 0.0;
    st->alpha = st->beta = st->pAp = 0.0;
    st->rel = 1.0;
    st->rel_true = 1.0;
  }
}

// The three partial sums of the PCG SpMV for row g with Ap_g = v.
template <typename T>
__device__ __forceinline__ void spmv_sums(const CgBufs<T>& b, int g, T v, double (&d)[3]) {
  const double av = static_cast<double>(v);
  const T dg = b.diag[g];
  const T zg = b.r[g] / dg;
  d[0] += static_cast<double>(b.p[g]) * av;
  d[1] += static_cast<double>(zg) * av;
  d[2] += av * av / static_cast<double>(dg);
}

// Block 0 passes the iteration count and the refresh step on to the vector
// kernel (see above).
__device__ __forceinline__ void spmv_iteration(PcgScalars* st, const SpmvWords& w) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->iter_cur = w.iter_done;
    st->spmv_mode = w.refresh_req;
  }
}

// XCD x of workgroup w (round-robin dispatch) and the number of workgroups
// of a grid of g (>= 8) on XCD x.
__device__ __forceinline__ int xcd_of_block() { return static_cast<int>(blockIdx.x) & 7; }
__device__ __forceinline__ int blocks_on_xcd(int x) {
  return (static_cast<int>(gridDim.x) - x + 7) >> 3;
}

// The slices a wave of the PCG SpMV handles: s0, s0 + stride, ... < hi, in
// the order BsrDev::xcd selects for value type T: XCD-contiguous ranges of
// ceil(nslices / 8) slices (grids of at least 8 workgroups, so every range
// has workgroups), or grid-stride over all slices.
struct SliceRange {
  int s0, stride, hi;
};
template <typename T>
__device__ __forceinline__ SliceRange slice_range(const BsrDev& A) {
  const int wpb = blockDim.x / kSliceRows, wave = threadIdx.x / kSliceRows;
  if (A.xcd[sizeof(T) == 4] && gridDim.x >= 8) {
    const int x = xcd_of_block(), per = (A.nslices + 7) >> 3;
    const int lo = min(A.nslices, x * per);
    return {lo + (static_cast<int>(blockIdx.x >> 3) * wpb + wave), blocks_on_xcd(x) * wpb,
            min(A.nslices, lo + per)};
  }
  return {static_cast<int>(blockIdx.x) * wpb + wave, static_cast<int>(gridDim.x) * wpb, A.nslices};
}

// Refresh step of the SpMV kernels (mixed precision): r64 = F - K64 u over
// the FP64 operator, |r64|^2 as partial slot 0 of partA, and the refresh
// count (block 0).
template <typename T>
__device__ __forceinline__ void refresh_done(const CgBufs<T>& b, PcgScalars* st, double (&v)[1],
                                             const CgRed& red, double* lds) {
  write_partials<1>(v, red.partA, lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->refreshes += 1;
}

// The SpMV's last workgroup reduces the partB buffer par (r.z, r.r, F.F of
// the current r) into sumB for the vector kernel that follows: the last
// workgroup has one slice per wave in either slice order, so it finishes its
// rows early and the reduction costs the SpMV nothing.
__device__ __forceinline__ void reduce_partB(PcgScalars* st, const CgRed& red, int par,
                                             double* lds) {
  if (blockIdx.x != gridDim.x - 1) return;
  const double* __restrict__ pb = red.partB + par * kPartB;
  double t[3] = {0.0, 0.0, 0.0};
  for (int q = threadIdx.x; q < red.nB; q += blockDim.x)
#pragma unroll
    for (int i = 0; i < 3; ++i) t[i] += pb[i * kMaxGrid + q];
  block_sum<3>(t, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < 3; ++i) st->sumB[i] = t[i];
}

template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_csr(CsrDev A, CgBufs<T> b, PcgScalars* st,
                                                        CgRed red, int par) {
  __shared__ double lds[kLds];
  const SpmvWords w = spmv_words(st);
  if (w.done) return;
  spmv_iteration(st, w);
  const int lg = threadIdx.x % G;
  const int stride = gridDim.x * (blockDim.x / G);
  if constexpr (sizeof(T) == 4) {
    if (w.refresh_req) {
      double v[1] = {0.0};
      for (int r = (blockIdx.x * blockDim.x + threadIdx.x) / G; r < A.n; r += stride) {
        const double s = csr_row<double, G>(A, A.v64, b.u64, r, lg);
        if (lg == 0) {
          const double ri = b.F[r] - s;
          b.r64[r] = ri;
          v[0] += ri * ri;
        }
      }
      refresh_done(b, st, v, red, lds);
      return;
    }
  }
  const T* val = vals<T>(A);
  double d[3] = {0.0, 0.0, 0.0};
  for (int r = (blockIdx.x * blockDim.x + threadIdx.x) / G; r < A.n; r += stride) {
    const T s = csr_row<T, G>(A, val, b.p, r, lg);
    if (lg == 0) {
      b.Ap[r] = s;
      spmv_sums(b, r, s, d);
    }
  }
  write_partials<3>(d, red.partA, lds);
  reduce_partB(st, red, par, lds);
}

// FP64 streams of the relative sliced ELL (REL) ask for 8 resident waves per
// SIMD with 3x3 blocks (58 VGPRs instead of 92: C2 83.3 -> 81.5 us per
// launch) and 7 with 6x6 blocks (66 VGPRs; 8 would spill). The absolute
// layout loads a block column per row and slot and needs the registers: under
// the same bound the FP64 solve of an operator with long rows (MPC tie of a
// face to its centre grid) took 137.8 ms instead of 91.5. The FP32 stream
// keeps its 88 VGPRs: asked for 6, 7 or 8 waves the compiler trims it to 56-58
// VGPRs, and the fewer loads in flight per wave cost more (34.5 -> 40.1-40.4 us).
template <typename T, int B, typename Acc = T, bool REL = false>
__global__ __launch_bounds__(kBlock, (sizeof(T) == 8 && REL) ? (B == 3 ? 8 : 7) : 1) void k_cg_spmv_bsr(
    BsrDev A, CgBufs<T> b, PcgScalars* st, CgRed red, int par) {
  __shared__ double lds[kLds];
  const int l = threadIdx.x & (kSliceRows - 1);
  const int nrows = A.nbr * B;
  const SliceRange sr = slice_range<T>(A);
  const int s0 = sr.s0;
  // the state words and the first slice's slot range load together (one
  // memory round trip before the stream starts, not two); the early exit
  // reads the slot range too, so its loads stay ahead of the branch
  const SpmvWords w = spmv_words(st);
  const Slots sl0 = s0 < sr.hi ? sell_slots<T>(A, s0) : Slots{0, 0};
  if (w.done) {
    if (sl0.km < 0) b.Ap[0] = T(0);  // never true
    return;
  }
  spmv_iteration(st, w);
  if constexpr (sizeof(T) == 4) {
    if (w.refresh_req) {
      double v[1] = {0.0};
      const SliceRange r64 = slice_range<double>(A);
      for (int s = r64.s0; s < r64.hi; s += r64.stride) {
        const double ku = sell_row<double, B>(A, A.s64, b.u64, s, l);
        const int g = kSliceRows * s + l;
        if (g < nrows) {
          const double ri = b.F[g] - ku;
          b.r64[g] = ri;
          v[0] += ri * ri;
        }
      }
      refresh_done(b, st, v, red, lds);
      return;
    }
  }
  const T* val = svals<T>(A);
  const bool split = (sizeof(T) == 4) && A.dblk != nullptr;
  double d[3] = {0.0, 0.0, 0.0};
  for (int s = s0; s < sr.hi; s += sr.stride) {
    const Acc acc = sell_row<T, B, Acc>(A, val, b.p, s, l, s == s0 ? sl0 : sell_slots<T>(A, s));
    const int g = kSliceRows * s + l;
    if (g < nrows) {
      // split FP32 stream: off-diagonal blocks summed in FP32, the corrected
      // FP64 diagonal block added in FP64
      const T v = split ? static_cast<T>(static_cast<double>(acc) +
                                         diag_block_row<T, B>(A.dblk, b.p, g))
                        : static_cast<T>(acc);
      b.Ap[g] = v;
      spmv_sums(b, g, v, d);
    }
  }
  write_partials<3>(d, red.partA, lds);
  reduce_partB(st, red, par, lds);
}

// The refresh step of the vector kernel: the decision on the FP64 true
// residual |r64|^2 = rr_true that the SpMV kernel just computed; on
// "continue", r = (T) r64 with the partials (r.z, r.r) of the new r, and the
// paused iteration marked resumed.
template <typename T>
__device__ __forceinline__ void cg_reload(int n, const CgBufs<T>& b, PcgScalars* st,
                                          const CgRed& red, const PcgParams& prm, double rr_true,
                                          const Head& h, int par, double* lds) {
  const double rel = sqrt(fabs(rr_true)) / st->norm_b;
  const bool lead = (blockIdx.x == 0 && threadIdx.x == 0);
  // stall history is double-buffered by refresh count (the SpMV kernel
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
      st->flags |= stop;
      st->done = 1;
    }
    return;
  }
  double v[2] = {0.0, 0.0};
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const T rk = static_cast<T>(b.r64[k]);
    const T zk = rk / b.diag[k];
    b.r[k] = rk;
    v[0] += static_cast<double>(rk) * static_cast<double>(zk);
    v[1] += static_cast<double>(rk) * static_cast<double>(rk);
  }
  write_partials<2>(v, red.partB + (par ^ 1) * kPartB, lds);
  if (lead) {
    st->resume_iter = h.iter_cur;
    st->refresh_req = 0;
  }
}

template <typename T, int NJA>
__global__ __launch_bounds__(kBlock) void k_cg_vector(int n, CgBufs<T> b, PcgScalars* st,
                                                      CgRed red, PcgParams prm, int mixed,
                                                      int par) {
  __shared__ double lds[kLds];
  __shared__ Head sh;
  const int n4 = n >> 2;
  const int tg = blockIdx.x * blockDim.x + threadIdx.x, ng = gridDim.x * blockDim.x;
  // row quads [q_lo, q_hi) of this workgroup's XCD range (CgBufs::xcd) or of
  // the whole vector, from q_lo + t0 in steps of nt
  int q_lo = 0, q_hi = n4, t0 = tg, nt = ng;
  if (b.xcd && gridDim.x >= 8) {
    const int per = (((n + kSliceRows - 1) / kSliceRows) + 7) >> 3;  // SpMV slices per XCD
    const int x = xcd_of_block();
    q_lo = min(n4, x * per * (kSliceRows / 4));
    q_hi = min(n4, (x + 1) * per * (kSliceRows / 4));
    t0 = static_cast<int>(blockIdx.x >> 3) * blockDim.x + threadIdx.x;
    nt = blocks_on_xcd(x) * blockDim.x;
  }
  HeadLoads<NJA> hl;
  head_issue<NJA>(st, red, hl);
  // this thread's first rows go out behind the head loads
  Quad<double> u0{};
  Quad<T> p0{}, ap0{}, d0{}, r0{};
  if (q_lo + t0 < q_hi) {
    const int k = 4 * (q_lo + t0);
    u0 = ld4(b.u64 + k);
    p0 = ld4(b.p + k);
    ap0 = ld4(b.Ap + k);
    d0 = ld4(b.diag + k);
    r0 = ld4(b.r + k);
  }
  double v6[6];
  Head h;
  head_finish<NJA>(hl, v6, h, lds, &sh);
  if (h.done) return;
  if constexpr (sizeof(T) == 4) {
    if (mixed && h.spmv_mode) {
      cg_reload<T>(n, b, st, red, prm, v6[0], h, par, lds);
      return;
    }
  }
  const int i = h.iter_cur;
  const bool resumed = (h.resume_iter == i);
  const bool lead = (blockIdx.x == 0 && threadIdx.x == 0);
  const double pAp = v6[0], zAp = v6[1], ApDAp = v6[2];
  const double rz = v6[3], rr = v6[4];
  double norm_b, rel;
  int stop = 0;
  bool pause = false;
  if (i == 0 && !resumed) {
    norm_b = sqrt(fabs(v6[5]));
    rel = 1.0;
    if (norm_b < 1e-300) {  // zero load: u = 0 after 0 iterations (impl:1822-1830)
      stop = kFlagZeroRhs;
      rel = 0.0;
    } else if (rz <= 0.0) {  // impl:1859-1862
      stop = kFlagRz0;
    } else if (prm.max_iters <= 0) {
      stop = kFlagMaxIter;
    }
  } else {
    norm_b = st->norm_b;
    rel = sqrt(fabs(rr)) / norm_b;
    if (!mixed) {
      if (rel < prm.tol) stop = kFlagConverged;
    } else if (!resumed &&
               (rel < prm.tol || rr < prm.refresh_delta * prm.refresh_delta * st->rr_refresh)) {
      pause = true;  // for an FP64 true-residual refresh
    }
    if (!stop && !pause && i >= prm.max_iters) stop = kFlagMaxIter;
  }
  if (!stop && !pause && !(pAp > 0.0)) stop = kFlagBreakdown;  // impl:691-696
  if (pause) {
    if (lead) {
      st->rz[0] = rz;
      st->rr = rr;
      st->pAp = pAp;
      st->rel = rel;
      st->refresh_req = 1;
    }
    return;
  }
  if (stop) {
    if (lead) {
      if (i == 0 && !resumed) st->norm_b = norm_b;
      st->rz[0] = rz;
      st->rr = rr;
      st->pAp = pAp;
      st->rel = (stop == kFlagBreakdown) ? 1e300 : rel;
      st->flags |= stop;
      st->done = 1;
    }
    return;
  }
  const double alpha = rz / pAp;
  const double rz_next = rz - 2.0 * alpha * zAp + alpha * alpha * ApDAp;
  const double beta = rz_next > 0.0 ? rz_next / rz : 0.0;
  const T a = static_cast<T>(alpha), bt = static_cast<T>(beta);
  double v[2] = {0.0, 0.0};
  auto rows4 = [&](int k, Quad<double>& u, Quad<T>& p, const Quad<T>& ap, const Quad<T>& d,
                   Quad<T>& r) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      u.v[j] += alpha * static_cast<double>(p.v[j]);
      r.v[j] = r.v[j] - a * ap.v[j];
      const T z = r.v[j] / d.v[j];
      p.v[j] = z + bt * p.v[j];
      v[0] += static_cast<double>(r.v[j]) * static_cast<double>(z);
      v[1] += static_cast<double>(r.v[j]) * static_cast<double>(r.v[j]);
    }
    st4(b.u64 + k, u);
    st4(b.r + k, r);
    st4(b.p + k, p);
  };
  if (q_lo + t0 < q_hi) rows4(4 * (q_lo + t0), u0, p0, ap0, d0, r0);
  for (int q = q_lo + t0 + nt; q < q_hi; q += nt) {
    const int k = 4 * q;
    Quad<double> u = ld4(b.u64 + k);
    Quad<T> p = ld4(b.p + k);
    const Quad<T> ap = ld4(b.Ap + k), d = ld4(b.diag + k);
    Quad<T> r = ld4(b.r + k);
    rows4(k, u, p, ap, d, r);
  }
  for (int k = 4 * n4 + tg; k < n; k += ng) {
    b.u64[k] += alpha * static_cast<double>(b.p[k]);
    const T r = b.r[k] - a * b.Ap[k];
    const T z = r / b.diag[k];
    b.r[k] = r;
    b.p[k] = z + bt * b.p[k];
    v[0] += static_cast<double>(r) * static_cast<double>(z);
    v[1] += static_cast<double>(r) * static_cast<double>(r);
  }
  write_partials<2>(v, red.partB + (par ^ 1) * kPartB, lds);
  if (lead) {
    if (i == 0 && !resumed) {
      st->norm_b = norm_b;
      st->rr_refresh = rr;
    }
    if (resumed) st->rr_refresh = rr;
    st->rz[0] = rz;
    st->rr = rr;
    st->rel = rel;
    st->pAp = pAp;
    st->alpha = alpha;
    st->beta = beta;
    st->iter_done = i + 1;
  }
}

// ---- long rows: the blocks past kSellCap -------------------------------------------------
// One workgroup per long block row (grid-stride over the rows): the row's
// remaining blocks (read from the BSR arrays, FP64 sums) are reduced over the
// workgroup and added to what the sliced-ELL pass wrote for the row's B
// scalar rows. MODE 0: y += rest. MODE 1 (PCG SpMV, x = p, y = Ap, z = r):
// y += rest, and the changes of the three SpMV sums (p.Ap, z.Ap, Ap.diag^-1.Ap)
// are added to partial slot blockIdx.x, which the SpMV pass wrote (grid <=
// its grid; one writer per slot, so the sums stay deterministic). MODE 2
// (refresh step of the PCG SpMV, FP64, x = u, y = r64): r64 -= rest and the
// change of r.r into slot 0 of the same partials.
// With the split FP32 stream the diagonal block is skipped (the SpMV added
// its FP64 form).
template <typename T, int B, int MODE>
__device__ __forceinline__ void overflow_rows(const BsrDev& A, const T* __restrict__ x,
                                              T* __restrict__ y, const T* __restrict__ z,
                                              const T* __restrict__ diag,
                                              double* __restrict__ part, double* lds) {
  const bool split = sizeof(T) == 4 && A.dblk != nullptr;
  const T* __restrict__ val = sizeof(T) == 4 ? reinterpret_cast<const T*>(A.v32)
                                             : reinterpret_cast<const T*>(A.v64);
  double corr[3] = {0.0, 0.0, 0.0};
  for (int li = blockIdx.x; li < A.nlong; li += gridDim.x) {
    const int I = A.lrows[li];
    const int o = A.brp[I], nb = A.brp[I + 1] - o;
    double acc[B];
#pragma unroll
    for (int r = 0; r < B; ++r) acc[r] = 0.0;
    // four blocks' loads in flight per thread (the sums keep their order)
#pragma unroll 4
    for (int k = kSellCap + threadIdx.x; k < nb; k += blockDim.x) {
      const int J = A.bci[o + k];
      if (split && J == I) continue;
      const T* vb = val + static_cast<size_t>(B * B) * o + k;
#pragma unroll
      for (int c = 0; c < B; ++c) {
        const double xc = static_cast<double>(x[B * J + c]);
#pragma unroll
        for (int r = 0; r < B; ++r) acc[r] += static_cast<double>(vb[static_cast<size_t>(r * B + c) * nb]) * xc;
      }
    }
    block_sum<B>(acc, lds);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int r = 0; r < B; ++r) {
        const int g = B * I + r;
        const double y0 = static_cast<double>(y[g]);
        if (MODE == 2) {
          const double rn = y0 - acc[r];
          y[g] = static_cast<T>(rn);
          corr[0] += rn * rn - y0 * y0;
        } else {
          const T yn = static_cast<T>(y0 + acc[r]);
          y[g] = yn;
          if (MODE == 1) {
            const double y1 = static_cast<double>(yn), dy = y1 - y0;
            corr[0] += static_cast<double>(x[g]) * dy;
            corr[1] += static_cast<double>(z[g] / diag[g]) * dy;
            corr[2] += (y1 * y1 - y0 * y0) / static_cast<double>(diag[g]);
          }
        }
      }
    }
  }
  if (threadIdx.x == 0) {
    if (MODE == 1)
#pragma unroll
      for (int i = 0; i < 3; ++i) part[i * kMaxGrid + blockIdx.x] += corr[i];
    else if (MODE == 2)
      part[blockIdx.x] += corr[0];
  }
}

// MODE 0: after a plain SpMV. MODE 1: after a PCG SpMV, which either formed
// Ap = K p (MODE 1 rows) or, in a mixed refresh step, r64 = F - K64 u (MODE 2
// rows on the FP64 blocks); it reads the same state words the SpMV read.
// 1,024 threads per long row: a face tied to one grid gives that grid's row
// thousands of blocks, a few per thread in one batch of loads (605k-DOF hex
// with its end face tied to the centre grid: mixed solve 58.4-59.4 ms with 256
// threads, 57.4-57.6 ms with 1,024; FP64 91.7 -> 90.3 ms)
constexpr int kOvfThreads = 1024;

template <typename T, int B, int MODE>
__global__ __launch_bounds__(kOvfThreads) void k_sell_overflow(BsrDev A, const T* __restrict__ x,
                                                          T* __restrict__ y,
                                                          const T* __restrict__ z,  // r (MODE 1)
                                                          const T* __restrict__ diag,
                                                          double* __restrict__ part,
                                                          const PcgScalars* st,
                                                          const double* __restrict__ u64,
                                                          double* __restrict__ r64) {
  __shared__ double lds[(kOvfThreads / 64) * 8];
  if (MODE == 1) {
    __shared__ int w[2];
    if (threadIdx.x == 0) {
      w[0] = st->done;
      w[1] = st->refresh_req;
    }
    __syncthreads();
    if (w[0]) return;
    if (sizeof(T) == 4 && w[1]) {
      overflow_rows<double, B, 2>(A, u64, r64, nullptr, nullptr, part, lds);
      return;
    }
  }
  overflow_rows<T, B, MODE>(A, x, y, z, diag, part, lds);
}

__global__ void k_cast_f64_f32(int n, const double* __restrict__ a, float* __restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    b[i] = static_cast<float>(a[i]);
}

// Streaming copy used to calibrate the HBM byte counters (known traffic).
__global__ __launch_bounds__(kBlock) void k_stream_copy(long long n, const double4* __restrict__ a,
                                                        double4* __restrict__ b) {
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    b[i] = a[i];
}

__global__ void k_fill_f64(long long n, double* __restrict__ a, double v) {
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    a[i] = v;
}

constexpr int kCsrG = 8;  // lanes per scalar CSR row

}  // namespace

// ---- launchers ------------------------------------------------------------------------
// The overflow pass of long rows after a sliced-ELL pass (see k_sell_overflow).
template <typename T, int MODE>
static void launch_overflow(const BsrDev& A, const T* x, T* y, const T* z, const T* diag,
                            double* part, int nprod, const PcgScalars* st, hipStream_t s,
                            const double* u64 = nullptr, double* r64 = nullptr) {
  if (A.nlong == 0) return;
  const int g = MODE == 0 ? std::min(A.nlong, kMaxGrid) : std::min(A.nlong, nprod);
  if (A.B == 3)
    hipLaunchKernelGGL((k_sell_overflow<T, 3, MODE>), dim3(g), dim3(kOvfThreads), 0, s, A, x, y, z, diag,
                       part, st, u64, r64);
  else
    hipLaunchKernelGGL((k_sell_overflow<T, 6, MODE>), dim3(g), dim3(kOvfThreads), 0, s, A, x, y, z, diag,
                       part, st, u64, r64);
}

template <typename T>
void launch_spmv_csr(const CsrDev& A, const T* x, T* y, hipStream_t s) {
  hipLaunchKernelGGL((k_spmv_csr<T, kCsrG>), dim3(grid_for(A.n, kBlock / kCsrG)), dim3(kBlock), 0,
                     s, A, x, y);
}

template <typename T>
void launch_spmv_bsr(const BsrDev& A, const T* x, T* y, hipStream_t s) {
  const dim3 g(grid_for(A.nslices, kBlock / kSliceRows));
  // the split FP32 stream sums in FP64 (cheap next to the matrix stream; an
  // FP32 sum gives back part of the iterations the split saves)
  const bool acc64 = sizeof(T) == 4 && A.dblk != nullptr;
  if (A.B == 3 && acc64)
    hipLaunchKernelGGL((k_spmv_bsr<T, 3, double>), g, dim3(kBlock), 0, s, A, x, y);
  else if (A.B == 3)
    hipLaunchKernelGGL((k_spmv_bsr<T, 3>), g, dim3(kBlock), 0, s, A, x, y);
  else if (acc64)
    hipLaunchKernelGGL((k_spmv_bsr<T, 6, double>), g, dim3(kBlock), 0, s, A, x, y);
  else
    hipLaunchKernelGGL((k_spmv_bsr<T, 6>), g, dim3(kBlock), 0, s, A, x, y);
  launch_overflow<T, 0>(A, x, y, nullptr, nullptr, nullptr, 0, nullptr, s);
}

template <typename T>
void launch_cg_init(int n, const CgBufs<T>& b, PcgScalars* st, const CgRed& red, hipStream_t s) {
  hipLaunchKernelGGL((k_cg_init<T>), dim3(red.nB), dim3(kBlock), 0, s, n, b, st, red);
}

// Launch with optional kernel timestamps: hipExtLaunchKernelGGL's start/stop
// events take the dispatch's own begin and end times (what a profiler reports),
// without the command-processor gap an event record adds around a launch.
#define VBH_LAUNCH(K, G, S, T0, T1, ...)                                        \
  do {                                                                          \
    if (T0) hipExtLaunchKernelGGL(K, G, dim3(kBlock), 0, S, T0, T1, 0u, __VA_ARGS__); \
    else hipLaunchKernelGGL(K, G, dim3(kBlock), 0, S, __VA_ARGS__);             \
  } while (0)

template <typename T>
void launch_cg_spmv(int format, const CsrDev& A, const BsrDev& Ab, const CgBufs<T>& b,
                    PcgScalars* st, const CgRed& red, hipStream_t s, int acc64, int position,
                    hipEvent_t t0, hipEvent_t t1) {
  if (sizeof(T) == 4 && Ab.dblk != nullptr) acc64 = 1;  // split FP32 stream: FP64 sums
  const dim3 g(red.nA);
  const int par = position & 1;
  if (format == 0)
    VBH_LAUNCH((k_cg_spmv_csr<T, kCsrG>), g, s, t0, t1, A, b, st, red, par);
  else if (format == 3 && acc64)
    VBH_LAUNCH((k_cg_spmv_bsr<T, 3, double>), g, s, t0, t1, Ab, b, st, red, par);
  else if (format == 3) {
    if constexpr (sizeof(T) == 8) {
      if (Ab.rk64) VBH_LAUNCH((k_cg_spmv_bsr<T, 3, T, true>), g, s, t0, t1, Ab, b, st, red, par);
      else VBH_LAUNCH((k_cg_spmv_bsr<T, 3>), g, s, t0, t1, Ab, b, st, red, par);
    } else {
      VBH_LAUNCH((k_cg_spmv_bsr<T, 3>), g, s, t0, t1, Ab, b, st, red, par);
    }
  } else if (acc64) {
    VBH_LAUNCH((k_cg_spmv_bsr<T, 6, double>), g, s, t0, t1, Ab, b, st, red, par);
  } else {
    if constexpr (sizeof(T) == 8) {
      if (Ab.rk64) VBH_LAUNCH((k_cg_spmv_bsr<T, 6, T, true>), g, s, t0, t1, Ab, b, st, red, par);
      else VBH_LAUNCH((k_cg_spmv_bsr<T, 6>), g, s, t0, t1, Ab, b, st, red, par);
    } else {
      VBH_LAUNCH((k_cg_spmv_bsr<T, 6>), g, s, t0, t1, Ab, b, st, red, par);
    }
  }
  if (format != 0)
    launch_overflow<T, 1>(Ab, b.p, b.Ap, b.r, b.diag, red.partA, red.nA, st, s, b.u64, b.r64);
}

// partial sums per thread for a producer grid of nprod blocks
static int head_nj(int nprod) {
  return nprod <= kBlock ? 1 : nprod <= 2 * kBlock ? 2 : nprod <= 4 * kBlock ? 4 : kMaxGrid / kBlock;
}
static_assert(kMaxGrid / kBlock == 8, "head_nj covers up to 8 partials per thread");

template <typename T>
void launch_cg_vector(int n, const CgBufs<T>& b, PcgScalars* st, const CgRed& red,
                      const PcgParams& prm, int mixed, int position, hipStream_t s,
                      hipEvent_t t0, hipEvent_t t1) {
  const dim3 g(red.nB);
  const int ja = head_nj(red.nA), par = position & 1;
#define VBH_VEC(JA)                                                                 \
  if (ja == JA) {                                                                   \
    VBH_LAUNCH((k_cg_vector<T, JA>), g, s, t0, t1, n, b, st, red, prm, mixed, par); \
    return;                                                                         \
  }
  VBH_VEC(1) VBH_VEC(2) VBH_VEC(4) VBH_VEC(8)
#undef VBH_VEC
}

#undef VBH_LAUNCH

int rows_per_block(int format) {
  return format == 0 ? kBlock / kCsrG : kBlock / kSliceRows;  // block formats: slices per block
}

void launch_cast_f64_to_f32(int n, const double* a, float* b, hipStream_t s) {
  hipLaunchKernelGGL(k_cast_f64_f32, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, n, a, b);
}

void launch_stream_copy(long long bytes, const void* a, void* b, hipStream_t s) {
  const long long n = bytes / static_cast<long long>(sizeof(double4));
  hipLaunchKernelGGL(k_stream_copy, dim3(4096), dim3(kBlock), 0, s, n,
                     static_cast<const double4*>(a), static_cast<double4*>(b));
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
                                     PcgScalars*, const CgRed&, hipStream_t, int, int, hipEvent_t,
                                     hipEvent_t);
template void launch_cg_spmv<float>(int, const CsrDev&, const BsrDev&, const CgBufs<float>&,
                                    PcgScalars*, const CgRed&, hipStream_t, int, int, hipEvent_t,
                                    hipEvent_t);
template void launch_cg_vector<double>(int, const CgBufs<double>&, PcgScalars*, const CgRed&,
                                       const PcgParams&, int, int, hipStream_t, hipEvent_t,
                                       hipEvent_t);
template void launch_cg_vector<float>(int, const CgBufs<float>&, PcgScalars*, const CgRed&,
                                      const PcgParams&, int, int, hipStream_t, hipEvent_t,
                                      hipEvent_t);

}  // namespace vbh
