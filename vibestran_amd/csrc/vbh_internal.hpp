// vibestran_amd/csrc/vbh_internal.hpp -- shared declarations of libvbh.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/vbh.h"
#include "vbh_error.hpp"

namespace vbh {

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
  void swap(DBuf& o) noexcept {
    T* tp = p;
    p = o.p;
    o.p = tp;
    const size_t tn = n;
    n = o.n;
    o.n = tn;
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
constexpr int kPartB = kMaxRed * kMaxGrid;  // stride of the two partB buffers

// Device-side PCG scalars (one instance per right-hand side). Written only by
// block 0 of a kernel, in fields no other block of that kernel reads.
struct PcgScalars {
  double rz[2];          // r.z of iteration i in slot i&1
  double alpha, beta, pAp, rr, rel, norm_b;
  double rr_refresh;     // recursive r.r at the last true-residual refresh (mixed)
  double rel_true;       // last FP64 true relative residual (mixed)
  double hist_rel[2];    // refresh history, slot = refresh count & 1
  int hist_stalls[2];
  int iter_done;         // completed iterations (written by cg_vector)
  int iter_cur;          // iteration in progress (written by cg_spmv)
  int resume_iter;       // iteration resumed after a refresh (written by cg_vector)
  int refresh_req;       // mixed: cg_vector paused for an FP64 refresh (read by cg_spmv)
  int spmv_mode;         // the last cg_spmv computed the refresh residual (read by cg_vector)
  int done, flags, refreshes;
  double sumB[3];        // r.z, r.r, F.F of the current r: the previous vector kernel's (or
                         // cg_init's) partials, reduced by cg_spmv's last workgroup
};

enum : int {
  kFlagBreakdown = 1,    // p.Ap <= 0 (cuda_pcg_backend_impl.hpp:691-696)
  kFlagConverged = 2,
  kFlagMaxIter = 4,
  kFlagRz0 = 8,          // r.z <= 0 at start (impl:1859-1862)
  kFlagZeroRhs = 16,     // ||F|| < 1e-300 (impl:1822-1830)
  kFlagRefresh = 32,     // block PCG, mixed: column paused for an FP64 true-residual refresh
  kFlagStall = 64        // mixed: true residual stopped improving
};

struct PcgParams {
  double tol;
  int max_iters;
  double refresh_delta;  // mixed: refresh when rr < delta^2 * rr_refresh
};

// Per-block partial sums passed from one kernel to the next. nA/nB are the
// grid sizes of the producing kernels (<= kMaxGrid).
struct CgRed {
  double* partA;  // [kMaxRed][kMaxGrid] from cg_spmv: p.Ap, z.Ap, Ap.diag^-1.Ap, or
                  // |F - K u|^2 (slot 0) when it computed a refresh residual
  double* partB;  // two buffers of kPartB: r.z, r.r, (F.F) from cg_init (buffer 0)
                  // and cg_vector (buffer (position + 1) & 1)
  int nA, nB;
};

template <typename T>
struct CgBufs {
  const double* F;
  const T* diag;
  T* r;
  T* p;
  T* Ap;
  double* u64;
  double* r64;  // mixed: FP64 true residual of a refresh (nullptr in FP64)
  // the vector kernel's rows follow the SpMV's XCD-contiguous slice ranges
  // (BsrDev::xcd of this value type), so each XCD updates the rows its SpMV
  // wrote and will gather
  int xcd = 0;
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
  // Sliced-ELL form of the same blocks, the layout the solver kernels read:
  // slice s = rows [64 s, 64 s + 64), one lane per scalar row. kmax_s =
  // koff[s+1] - koff[s] block slots per row (padded with zero blocks); the B
  // values of row g's k-th block are at ((koff[s] + k) * 64 + (g - 64 s)) * B,
  // and its block column at sidx[(koff[s] + k) * sell_nps(B) + node - node0(s)]
  // with node = g / B, node0(s) = 64 s / B.
  int nslices = 0;
  const int* koff = nullptr;
  const int* sidx = nullptr;
  // The same block columns as 16-bit offsets from node0(s), when every one
  // fits (banded numberings: a structured hex's neighbours lie within
  // +-(ny+1)(nz+1) nodes); the SpMVs then read these instead of sidx.
  const short* sidx16 = nullptr;
  const double* s64 = nullptr;
  const float* s32 = nullptr;
  // Row-sum preserving FP32 stream: when dblk is set, s32 holds zeros in
  // every diagonal node block and dblk (FP64, B*B per block row, row-major)
  // holds that block corrected so each scalar row's per-component sums equal
  // the FP64 operator's; FP32 SpMVs sum the off-diagonal blocks in FP32 and
  // add the diagonal block's product in FP64.
  const double* dblk = nullptr;
  // Long block rows (more than kSellCap blocks, e.g. an MPC master coupled to
  // a whole face): the sliced ELL holds their first kSellCap blocks only, so
  // one row does not widen its slice, and a separate pass adds the rest
  // (blocks kSellCap.. of the row in the BSR arrays above) after the SpMV.
  int nlong = 0;
  const int* lrows = nullptr;  // block rows with more than kSellCap blocks
  // Slice order of the PCG SpMV per value type (0 FP64 stream, 1 FP32):
  // 1 = XCD-contiguous (workgroup w runs on XCD w mod 8, so the slices are
  // cut into 8 contiguous ranges, one per XCD, and the gathered vector of
  // each range is fetched into that XCD's L2 only; grids of at least 8
  // workgroups), 0 = grid-stride over all.
  int xcd[2] = {0, 0};
  // Relative sliced ELL (set when it streams fewer bytes; replaces koff /
  // sidx / s64 / s32 for the SpMVs): slice s holds the union of its nodes'
  // block-column offsets J - I, rel_off[rel_koff[s] .. rel_koff[s+1]), one
  // uniform offset per slot for the whole slice, so no per-row column index is
  // read; a row's value block is zero where its node has no such neighbour.
  // The FP32 list leaves out offset 0 when dblk carries the diagonal blocks.
  const int* rk64 = nullptr;  // slot offsets per slice (nslices + 1), FP64 list
  const int* ro64 = nullptr;  // block-column offsets, FP64 list
  const double* rv64 = nullptr;
  const int* rk32 = nullptr;
  const int* ro32 = nullptr;
  const float* rv32 = nullptr;
};

constexpr int kSliceRows = 64;
constexpr int kSellCap = 128;  // block slots per row kept in the sliced ELL
__host__ __device__ constexpr int sell_nps(int B) { return kSliceRows / B + 2; }

// ---- kernel launchers (vbh_kernels.hip) ------------------------------------------
int grid_for(long long work_items, int items_per_block);
// Grid of the PCG vector kernels: four rows per thread, at most two
// workgroups per compute unit. Every workgroup reduces all partial sums of
// the previous kernels, so fewer, fuller workgroups cost less than one quad
// per thread over more of them (MI355X, 605k rows: 12.7 us at 512
// workgroups, 14.7 us at 591).
inline int vector_grid(long long n, int n_cu) {
  const int g = grid_for(n, kBlock * 4);
  return g < 2 * n_cu ? g : 2 * n_cu;
}

template <typename T>
void launch_spmv_csr(const CsrDev& A, const T* x, T* y, hipStream_t s);
template <typename T>
void launch_spmv_bsr(const BsrDev& A, const T* x, T* y, hipStream_t s);

// Jacobi-PCG building blocks (vector type T: double or float; dots in double).
int rows_per_block(int format);
template <typename T>
void launch_cg_init(int n, const CgBufs<T>& b, PcgScalars* st, const CgRed& red, hipStream_t s);
// acc64: accumulate the SpMV rows in FP64 (products of the stored values).
// t0/t1 (optional): events that receive the kernel's own start and end times.
// position: as for launch_cg_vector (the SpMV reduces the partB buffer that
// the vector kernel at the same position would read).
template <typename T>
void launch_cg_spmv(int format, const CsrDev& A, const BsrDev& Ab, const CgBufs<T>& b,
                    PcgScalars* st, const CgRed& red, hipStream_t s, int acc64, int position,
                    hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
// position: the launch's place in its sequence of iterations (a captured
// graph or a run of direct launches); sequences have an even length and
// start after cg_init (which writes partB buffer 0).
template <typename T>
void launch_cg_vector(int n, const CgBufs<T>& b, PcgScalars* st, const CgRed& red,
                      const PcgParams& prm, int mixed, int position, hipStream_t s,
                      hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);

// Device address of the allocation-free solver's scratch array on `device`
// (current), taken once per device (vbh_noalloc.hip).
void scratch_init(int device);

// Builds the sliced-ELL arrays of a device BSR (vbh_sell.hip).
struct Ctx;
void build_sell(Ctx& c);
// Sorted distinct block-column offsets (bci - I) of the block rows each slice
// overlaps, as a CSR over slices (vbh_assemble.hip).
void sell_slice_offsets(const BsrDev& A, int nslices, std::vector<int>& koff,
                        std::vector<int>& offs, hipStream_t s);

// Plain elementwise helpers.
void launch_cast_f64_to_f32(int n, const double* a, float* b, hipStream_t s);
void launch_fill_f64(long long n, double* a, double v, hipStream_t s);
void launch_stream_copy(long long bytes, const void* a, void* b, hipStream_t s);

}  // namespace vbh
