// vibestran_amd/csrc/vbh_sell.hip -- sliced-ELL layout of the block operator.
//
// The solver's SpMV assigns one lane to each scalar row and walks the row's
// blocks; with 64-row slices the B values a lane needs per block are
// contiguous and the lanes of a wave read one contiguous span per block slot,
// with no cross-lane reduction. Built on the device from the BSR operator.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "vbh_ctx.hpp"

namespace vbh {
namespace {

__global__ void k_sell_kmax(int nslices, int nbr, int B, const int* __restrict__ brp,
                            int* __restrict__ kcnt) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslices) return;
  const int n0 = (kSliceRows * s) / B;
  int n1 = (kSliceRows * s + kSliceRows - 1) / B;
  if (n1 > nbr - 1) n1 = nbr - 1;
  int km = 0;
  for (int I = n0; I <= n1; ++I) {
    const int nb = min(brp[I + 1] - brp[I], kSellCap);  // longer rows: the rest overflows
    km = nb > km ? nb : km;
  }
  kcnt[s] = km;
}

template <int B>
__global__ void k_sell_fill(int nslices, int nbr, const int* __restrict__ brp,
                            const int* __restrict__ bci, const double* __restrict__ bv,
                            const int* __restrict__ koff, double* __restrict__ sv,
                            int* __restrict__ sidx) {
  const int s = blockIdx.x;
  const int l = threadIdx.x;  // 0..63
  if (s >= nslices) return;
  const int k0 = koff[s], km = koff[s + 1] - k0;
  const int g = kSliceRows * s + l;
  const int node0 = (kSliceRows * s) / B;
  if (g < nbr * B) {
    const int I = g / B, r = g % B;
    const int o = brp[I], nb = brp[I + 1] - o;
    for (int k = 0; k < km; ++k)
      for (int c = 0; c < B; ++c)
        sv[(static_cast<size_t>(k0 + k) * kSliceRows + l) * B + c] =
            (k < nb) ? bv[static_cast<size_t>(B * B) * o + static_cast<size_t>(r * B + c) * nb + k] : 0.0;
  } else {
    for (int k = 0; k < km; ++k)
      for (int c = 0; c < B; ++c) sv[(static_cast<size_t>(k0 + k) * kSliceRows + l) * B + c] = 0.0;
  }
  constexpr int NPS = sell_nps(B);
  if (l < NPS) {
    const int I = node0 + l;
    const int o = (I < nbr) ? brp[I] : 0;
    const int nb = (I < nbr) ? brp[I + 1] - o : 0;
    for (int k = 0; k < km; ++k)
      sidx[static_cast<size_t>(k0 + k) * NPS + l] = (k < nb) ? bci[o + k] : (I < nbr ? I : 0);
  }
}

// 16-bit form of the slot block columns: J - node0(s); entries of rows past
// the last node (padding, zero values) become 0. *bad = 1 when an offset does
// not fit.
template <int B>
__global__ void k_sell_idx16(int nslices, int nbr, const int* __restrict__ koff,
                             const int* __restrict__ sidx, short* __restrict__ idx16,
                             int* __restrict__ bad) {
  const int s = blockIdx.x;
  const int l = threadIdx.x;
  constexpr int NPS = sell_nps(B);
  if (s >= nslices || l >= NPS) return;
  const int node0 = (kSliceRows * s) / B;
  const int I = node0 + l;
  for (int k = koff[s]; k < koff[s + 1]; ++k) {
    const size_t at = static_cast<size_t>(k) * NPS + l;
    int d = (I < nbr) ? sidx[at] - node0 : 0;
    if (d < -32768 || d > 32767) {
      *bad = 1;
      d = 0;
    }
    idx16[at] = static_cast<short>(d);
  }
}

// Diagonal node block of block row I in FP64 plus the symmetrised correction
// (e + e^T)/2, e[r][c] = sum over the off-diagonal blocks of the row of
// (v - (double)(float)v) at entry (r, c): with FP32 off-diagonal blocks and
// this block, the per-component sums of every scalar row equal the FP64
// operator's (up to the symmetrisation), so fields that are locally constant
// per component -- the smooth vectors late PCG iterations consist of -- are
// multiplied without the FP32 rounding of the blocks around them.
template <int B>
__global__ void k_diag_rowsum(int nbr, const int* __restrict__ brp, const int* __restrict__ bci,
                              const double* __restrict__ bv, double* __restrict__ dblk,
                              int* __restrict__ kdiag) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nbr) return;
  const int o = brp[I], nb = brp[I + 1] - o;
  double e[B][B], a[B][B];
  for (int r = 0; r < B; ++r)
    for (int c = 0; c < B; ++c) e[r][c] = a[r][c] = 0.0;
  int kd = -1;
  for (int q = 0; q < nb; ++q) {
    const bool diag = bci[o + q] == I;
    if (diag) kd = q;
    for (int r = 0; r < B; ++r)
      for (int c = 0; c < B; ++c) {
        const double x = bv[static_cast<size_t>(B * B) * o + static_cast<size_t>(r * B + c) * nb + q];
        if (diag) a[r][c] = x;
        else e[r][c] += x - static_cast<double>(static_cast<float>(x));
      }
  }
  for (int r = 0; r < B; ++r)
    for (int c = 0; c < B; ++c)
      dblk[(static_cast<size_t>(I) * B + r) * B + c] = a[r][c] + 0.5 * (e[r][c] + e[c][r]);
  kdiag[I] = kd;
}

// Zero the diagonal node block of every row in the FP32 sliced-ELL values.
template <int B>
__global__ void k_sell_zero_diag(int nslices, int nbr, const int* __restrict__ koff,
                                 const int* __restrict__ kdiag, float* __restrict__ s32) {
  const int s = blockIdx.x, l = threadIdx.x;
  if (s >= nslices) return;
  const int g = kSliceRows * s + l;
  if (g >= nbr * B) return;
  const int kd = kdiag[g / B];
  if (kd < 0 || kd >= koff[s + 1] - koff[s]) return;  // none, or in a long row's overflow
  for (int c = 0; c < B; ++c) s32[(static_cast<size_t>(koff[s] + kd) * kSliceRows + l) * B + c] = 0.0f;
}

// Values of the relative sliced ELL: slice s, lane l = row 64 s + l of node
// I; slot k holds the row's block at column I + off[k] (found by binary
// search in the sorted BSR row) or zeros.
template <int B, typename T>
__global__ void k_rel_fill(int nslices, int nbr, const int* __restrict__ brp,
                           const int* __restrict__ bci, const double* __restrict__ bv,
                           const int* __restrict__ rk, const int* __restrict__ ro,
                           T* __restrict__ rv) {
  const int s = blockIdx.x;
  const int l = threadIdx.x;  // 0..63
  if (s >= nslices) return;
  const int k0 = rk[s], km = rk[s + 1] - k0;
  const int g = kSliceRows * s + l;
  const int I = g / B, r = g - I * B;
  const bool row = g < nbr * B;
  const int o = row ? brp[I] : 0, nb = row ? brp[I + 1] - o : 0;
  for (int k = 0; k < km; ++k) {
    const int J = I + ro[k0 + k];
    int kk = -1;
    if (row) {
      int lo = 0, hi = nb - 1;
      while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int c = bci[o + mid];
        if (c == J) {
          kk = mid;
          break;
        }
        if (c < J) lo = mid + 1;
        else hi = mid - 1;
      }
    }
    T* dst = rv + (static_cast<size_t>(k0 + k) * kSliceRows + l) * B;
    for (int c = 0; c < B; ++c)
      dst[c] = kk >= 0 ? static_cast<T>(bv[static_cast<size_t>(B * B) * o +
                                          static_cast<size_t>(r * B + c) * nb + kk])
                       : T(0);
  }
}

__global__ void k_to_f32(long long m, const double* __restrict__ a, float* __restrict__ b) {
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < m;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    b[i] = static_cast<float>(a[i]);
}

}  // namespace

bool build_sell_relative(Ctx& c);

// Fills the absolute sliced ELL of ktot slots: FP64 and FP32 values, block
// columns (16-bit where every offset fits), and the FP32 diagonal zeroed
// where the corrected FP64 diagonal blocks carry it.
static void build_sell_absolute(Ctx& c, int ktot) {
  hipStream_t s = c.stream;
  BsrDev& A = c.bsrd;
  const int B = A.B, nslices = A.nslices;
  const size_t nval = static_cast<size_t>(ktot) * kSliceRows * B;
  c.sell_v64.resize(nval > 0 ? nval : 1);
  c.sell_v32.resize(nval > 0 ? nval : 1);
  c.sell_idx.resize(static_cast<size_t>(ktot) * sell_nps(B) + 1);
  if (B == 3)
    hipLaunchKernelGGL(k_sell_fill<3>, dim3(nslices), dim3(kSliceRows), 0, s, nslices, A.nbr, A.brp,
                       A.bci, A.v64, c.sell_koff.p, c.sell_v64.p, c.sell_idx.p);
  else
    hipLaunchKernelGGL(k_sell_fill<6>, dim3(nslices), dim3(kSliceRows), 0, s, nslices, A.nbr, A.brp,
                       A.bci, A.v64, c.sell_koff.p, c.sell_v64.p, c.sell_idx.p);
  hipLaunchKernelGGL(k_to_f32, dim3(grid_for(static_cast<long long>(nval), 1024)), dim3(256), 0, s,
                     static_cast<long long>(nval), c.sell_v64.p, c.sell_v32.p);
  if (A.dblk) {
    if (B == 3)
      hipLaunchKernelGGL(k_sell_zero_diag<3>, dim3(nslices), dim3(kSliceRows), 0, s, nslices, A.nbr,
                         c.sell_koff.p, c.sell_kdiag.p, c.sell_v32.p);
    else
      hipLaunchKernelGGL(k_sell_zero_diag<6>, dim3(nslices), dim3(kSliceRows), 0, s, nslices, A.nbr,
                         c.sell_koff.p, c.sell_kdiag.p, c.sell_v32.p);
  }
  // 16-bit block columns when every offset fits (VBH_SELL_IDX16=0 keeps the
  // 32-bit ones)
  A.sidx16 = nullptr;
  c.sell_idx16.release();
  const char* i16_env = std::getenv("VBH_SELL_IDX16");
  if (!(i16_env && i16_env[0] == '0')) {
    c.sell_idx16.resize(static_cast<size_t>(ktot) * sell_nps(B) + 1);
    DBuf<int> bad;
    bad.resize(1);
    VBH_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    if (B == 3)
      hipLaunchKernelGGL(k_sell_idx16<3>, dim3(nslices), dim3(kSliceRows), 0, s, nslices, A.nbr,
                         c.sell_koff.p, c.sell_idx.p, c.sell_idx16.p, bad.p);
    else
      hipLaunchKernelGGL(k_sell_idx16<6>, dim3(nslices), dim3(kSliceRows), 0, s, nslices, A.nbr,
                         c.sell_koff.p, c.sell_idx.p, c.sell_idx16.p, bad.p);
    int h_bad = 0;
    VBH_HIP(hipMemcpyAsync(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    VBH_HIP(hipStreamSynchronize(s));
    if (h_bad) c.sell_idx16.release();
    else A.sidx16 = c.sell_idx16.p;
  }
  VBH_HIP(hipStreamSynchronize(s));
  A.sidx = c.sell_idx.p;
  A.s64 = c.sell_v64.p;
  A.s32 = c.sell_v32.p;
}

void build_sell(Ctx& c) {
  hipStream_t s = c.stream;
  BsrDev& A = c.bsrd;
  const int B = A.B;
  const int nrows = A.nbr * B;
  const int nslices = (nrows + kSliceRows - 1) / kSliceRows;
  DBuf<int> kcnt;
  kcnt.resize(static_cast<size_t>(nslices) + 1);
  c.sell_koff.resize(static_cast<size_t>(nslices) + 1);
  hipLaunchKernelGGL(k_sell_kmax, dim3((nslices + 255) / 256), dim3(256), 0, s, nslices, A.nbr, B,
                     A.brp, kcnt.p);
  // exclusive scan of kcnt (length nslices+1; the last slot is a zero pad)
  VBH_HIP(hipMemsetAsync(kcnt.p + nslices, 0, sizeof(int), s));
  {
    size_t tmp = 0;
    VBH_HIP(rocprim::exclusive_scan(nullptr, tmp, kcnt.p, c.sell_koff.p, 0,
                                    static_cast<size_t>(nslices) + 1, rocprim::plus<int>(), s));
    DBuf<char> t;
    t.resize(tmp > 0 ? tmp : 1);
    VBH_HIP(rocprim::exclusive_scan(t.p, tmp, kcnt.p, c.sell_koff.p, 0,
                                    static_cast<size_t>(nslices) + 1, rocprim::plus<int>(), s));
  }
  int ktot = 0;
  VBH_HIP(hipMemcpyAsync(&ktot, c.sell_koff.p + nslices, sizeof(int), hipMemcpyDeviceToHost, s));
  VBH_HIP(hipStreamSynchronize(s));
  A.dblk = nullptr;
  if (c.rowsum_fp32) {
    c.sell_dblk.resize(static_cast<size_t>(A.nbr) * B * B);
    c.sell_kdiag.resize(static_cast<size_t>(A.nbr));
    if (B == 3)
      hipLaunchKernelGGL(k_diag_rowsum<3>, dim3((A.nbr + 255) / 256), dim3(256), 0, s, A.nbr, A.brp,
                         A.bci, A.v64, c.sell_dblk.p, c.sell_kdiag.p);
    else
      hipLaunchKernelGGL(k_diag_rowsum<6>, dim3((A.nbr + 255) / 256), dim3(256), 0, s, A.nbr, A.brp,
                         A.bci, A.v64, c.sell_dblk.p, c.sell_kdiag.p);
    A.dblk = c.sell_dblk.p;
  } else {
    c.sell_dblk.release();
    c.sell_kdiag.release();
  }
  // long rows: their blocks past kSellCap are added by the overflow pass
  {
    std::vector<int> brp(static_cast<size_t>(A.nbr) + 1), lr;
    VBH_HIP(hipMemcpyAsync(brp.data(), A.brp, sizeof(int) * brp.size(), hipMemcpyDeviceToHost, s));
    VBH_HIP(hipStreamSynchronize(s));
    for (int I = 0; I < A.nbr; ++I)
      if (brp[static_cast<size_t>(I) + 1] - brp[static_cast<size_t>(I)] > kSellCap) lr.push_back(I);
    c.sell_lrows.upload(lr.data(), lr.size(), s);
    A.nlong = static_cast<int>(lr.size());
    A.lrows = lr.empty() ? nullptr : c.sell_lrows.p;
  }
  A.nslices = nslices;
  A.koff = c.sell_koff.p;
  c.sell_padded_blocks = static_cast<long long>(ktot) * kSliceRows / B;
  c.sell_slots = ktot;
  // the absolute layout is filled only where the relative one is not taken
  // (no reader of the absolute arrays runs while BsrDev::rk64 is set)
  A.sidx = nullptr;
  A.sidx16 = nullptr;
  A.s64 = nullptr;
  A.s32 = nullptr;
  if (build_sell_relative(c)) {
    c.sell_v64.release();
    c.sell_v32.release();
    c.sell_idx.release();
    c.sell_idx16.release();
  } else {
    build_sell_absolute(c, ktot);
  }
  // PCG SpMV slice order (BsrDev::xcd): XCD-contiguous ranges for the FP32
  // stream of 3x3 blocks (C2 hex: 35.4 -> 33.7 us per launch); the FP64
  // stream (81.6 -> 82.6 us) and 6x6 shell blocks are faster grid-stride
  // (VBH_SPMV_XCD=0: grid-stride for both, a diagnosis knob)
  const char* xe = std::getenv("VBH_SPMV_XCD");
  A.xcd[0] = 0;
  A.xcd[1] = (B == 3 && !(xe && xe[0] == '0')) ? 1 : 0;
}

// The relative sliced ELL (BsrDev::rk64 ...), used when its slot count stays
// within 10 % of the absolute one's and no row overflows: on meshes numbered
// by a structured sweep every slice sees the same few neighbour offsets, so
// the per-row block columns (and, for the FP32 stream, the zeroed diagonal
// blocks) drop out of the operator stream. VBH_SELL_REL=0 keeps the absolute
// layout.
bool build_sell_relative(Ctx& c) {
  BsrDev& A = c.bsrd;
  A.rk64 = A.ro64 = A.rk32 = A.ro32 = nullptr;
  A.rv64 = nullptr;
  A.rv32 = nullptr;
  c.rel_k64.release();
  c.rel_o64.release();
  c.rel_k32.release();
  c.rel_o32.release();
  c.rel_v64.release();
  c.rel_v32.release();
  c.rel_slots64 = c.rel_slots32 = 0;
  const char* env = std::getenv("VBH_SELL_REL");
  if ((env && env[0] == '0') || A.nlong > 0) return false;
  hipStream_t s = c.stream;
  const int B = A.B, nbr = A.nbr, nslices = A.nslices;
  std::vector<int> koff, offs;
  sell_slice_offsets(A, nslices, koff, offs, s);
  const bool split = A.dblk != nullptr;  // FP32 list without the diagonal offset
  std::vector<int> k64(static_cast<size_t>(nslices) + 1, 0), k32(static_cast<size_t>(nslices) + 1, 0);
  std::vector<int> o64, o32;
  o64.reserve(offs.size());
  o32.reserve(offs.size());
  for (int sl = 0; sl < nslices; ++sl) {
    const int a = koff[static_cast<size_t>(sl)], b = koff[static_cast<size_t>(sl) + 1];
    if (b - a > kSellCap) return false;
    for (int q = a; q < b; ++q) {
      const int off = offs[static_cast<size_t>(q)];
      o64.push_back(off);
      if (!(split && off == 0)) o32.push_back(off);
    }
    k64[static_cast<size_t>(sl) + 1] = static_cast<int>(o64.size());
    k32[static_cast<size_t>(sl) + 1] = static_cast<int>(o32.size());
  }
  if (static_cast<double>(o64.size()) > 1.1 * static_cast<double>(c.sell_slots)) return false;
  c.rel_k64.upload(k64.data(), k64.size(), s);
  c.rel_o64.upload(o64.data(), o64.size(), s);
  c.rel_k32.upload(k32.data(), k32.size(), s);
  c.rel_o32.upload(o32.data(), std::max<size_t>(o32.size(), 1), s);
  c.rel_v64.resize(o64.size() * kSliceRows * B);
  c.rel_v32.resize(std::max<size_t>(o32.size(), 1) * kSliceRows * B);
  if (B == 3) {
    hipLaunchKernelGGL((k_rel_fill<3, double>), dim3(nslices), dim3(kSliceRows), 0, s, nslices, nbr,
                       A.brp, A.bci, A.v64, c.rel_k64.p, c.rel_o64.p, c.rel_v64.p);
    hipLaunchKernelGGL((k_rel_fill<3, float>), dim3(nslices), dim3(kSliceRows), 0, s, nslices, nbr,
                       A.brp, A.bci, A.v64, c.rel_k32.p, c.rel_o32.p, c.rel_v32.p);
  } else {
    hipLaunchKernelGGL((k_rel_fill<6, double>), dim3(nslices), dim3(kSliceRows), 0, s, nslices, nbr,
                       A.brp, A.bci, A.v64, c.rel_k64.p, c.rel_o64.p, c.rel_v64.p);
    hipLaunchKernelGGL((k_rel_fill<6, float>), dim3(nslices), dim3(kSliceRows), 0, s, nslices, nbr,
                       A.brp, A.bci, A.v64, c.rel_k32.p, c.rel_o32.p, c.rel_v32.p);
  }
  VBH_HIP(hipStreamSynchronize(s));
  A.rk64 = c.rel_k64.p;
  A.ro64 = c.rel_o64.p;
  A.rv64 = c.rel_v64.p;
  A.rk32 = c.rel_k32.p;
  A.ro32 = c.rel_o32.p;
  A.rv32 = c.rel_v32.p;
  c.rel_slots64 = static_cast<long long>(o64.size());
  c.rel_slots32 = static_cast<long long>(o32.size());
  return true;
}

}  // namespace vbh
