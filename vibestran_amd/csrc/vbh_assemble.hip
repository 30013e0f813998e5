// vibestran_amd/csrc/vbh_assemble.hip -- DOF numbering, batched element
// stiffness and lower-CSR assembly on the device (gfx950).
//
// Replaces the reference's per-element host loop
//   make_element -> global_dof_indices -> stiffness_matrix -> to_row_major ->
//   MpcHandler fast path -> SparseMatrixBuilder::add_element_stiffness ->
//   build_csr
// (reference src/solver/assembly_parallel.hpp:68-110, src/core/sparse_matrix.cpp:16-150)
// with four device passes:
//   1. element DOF lists           (solid_global_dof_indices, solid_elements.cpp:29-42)
//   2. topological lower pattern   every (max(gi,gj), min(gi,gj)) pair of every
//                                  element, radix-sorted and de-duplicated;
//                                  when every node carries the same DOFs (3
//                                  for solids, 6 for CQUAD4) the node pairs are
//                                  sorted instead (36 keys per CHEXA8, not 300)
//                                  and each is expanded into its DOF slots in
//                                  CSR order (pass 2': same slots, same sums in
//                                  the same order)
//   3. element matrices           one wave64 per element; each lane owns a few
//                                  lower entries, accumulated with the closed
//                                  isotropic form lam*a_p*b_q + mu*a_q*b_p +
//                                  mu*delta_pq*(a.b) (== B^T D B for the
//                                  isotropic D of solid_elements.cpp:17-27);
//                                  EAS modes condensed through a 9x9 Cholesky in
//                                  LDS (solid_elements.cpp:1201-1204). Every
//                                  entry is stored at its position in the pair
//                                  list of pass 2 (coalesced, no atomics).
//   3b. slot sums                  the pair keys were sorted together with
//                                  their positions (stable radix sort), so each
//                                  CSR slot's contributions form one run in
//                                  element order; one thread per slot sums the
//                                  run's nonzero entries in that order. Entries
//                                  that are exactly 0.0 are not added, and a
//                                  slot without any nonzero is marked empty
//                                  (sparse_matrix.cpp:30-33).
//   4. compaction                  empty slots vanish, so the pattern equals
//                                  the reference's.
// K is therefore bit-identical from run to run: every slot is summed in
// element order, one fixed order of the serial assembly (the reference's TBB
// merge order is thread-dependent, assembly_parallel.hpp:84-100), the order
// of the CPU oracle's stable build_csr.
#include <cstdlib>
#include <cstring>
#include <numeric>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <vector>

#include "vbh_ctx.hpp"
#include "vbh_elem.hpp"

namespace vbh {
namespace {

// ---- lower-triangle pair tables ---------------------------------------------------
__constant__ PairTable<24> c_pairs24;
__constant__ PairTable<18> c_pairs18;
__constant__ PairTable<12> c_pairs12;
__constant__ PairTable<9> c_pairs9;

template <int N>
__device__ __forceinline__ unsigned short pair_at(int t);
template <>
__device__ __forceinline__ unsigned short pair_at<24>(int t) { return c_pairs24.v[t]; }
template <>
__device__ __forceinline__ unsigned short pair_at<18>(int t) { return c_pairs18.v[t]; }
template <>
__device__ __forceinline__ unsigned short pair_at<12>(int t) { return c_pairs12.v[t]; }
template <>
__device__ __forceinline__ unsigned short pair_at<9>(int t) { return c_pairs9.v[t]; }

enum ElemKind : int { kHexEas = 0, kHexSri = 1, kPentaEas = 2, kTet4 = 3 };

template <int K>
struct Kind {
  static constexpr int NN = (K == kTet4) ? 4 : (K == kPentaEas ? 6 : 8);
  static constexpr int ND = 3 * NN;
  static constexpr int NL = ND * (ND + 1) / 2;
  static constexpr bool EAS = (K == kHexEas || K == kPentaEas);
  static constexpr int SU = (NL + 63) / 64;
  static constexpr int SA = (ND * 9 + 63) / 64;
};


// ---- 1. element DOF lists ------------------------------------------------------------
__global__ void k_gdofs(int nelem, const int* __restrict__ etype, const int* __restrict__ eptr,
                        const int* __restrict__ conn, const int* __restrict__ eq,
                        int* __restrict__ gd) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const int t = etype[e];
  const int o = eptr[e], nn = eptr[e + 1] - o;
  int* g = gd + static_cast<size_t>(e) * kGdStride;
  int k = 0;
  if (t == VBH_CQUAD4) {
    for (int a = 0; a < nn; ++a)
      for (int d = 0; d < 6; ++d) g[k++] = eq[6 * conn[o + a] + d];
  } else {
    for (int a = 0; a < nn; ++a)
      for (int d = 0; d < 3; ++d) g[k++] = eq[6 * conn[o + a] + d];
  }
  for (; k < kGdStride; ++k) g[k] = -1;
}

// ---- 2. topological pattern keys ----------------------------------------------------------
// keys[idx] = (row << bits) | col of the element's lower pair, src[idx] =
// off + idx: the pair's position in the concatenated list of all kinds.
// With eid, also the element of each pair (meshes of several element kinds,
// whose slots are summed in element order across the kind groups).
template <int ND>
__global__ void k_pair_keys(int nsel, const int* __restrict__ sel, const int* __restrict__ gd,
                            unsigned long long* __restrict__ keys, int bits,
                            unsigned* __restrict__ src = nullptr, unsigned off = 0,
                            int* __restrict__ eid = nullptr) {
  constexpr int NL = ND * (ND + 1) / 2;
  const long long idx = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (idx >= static_cast<long long>(nsel) * NL) return;
  if (src) src[idx] = off + static_cast<unsigned>(idx);
  if (eid) eid[idx] = sel[idx / NL];
  const int s = static_cast<int>(idx / NL), t = static_cast<int>(idx % NL);
  const unsigned short pr = pair_at<ND>(t);
  const int e = sel[s];
  const int gi = gd[static_cast<size_t>(e) * kGdStride + (pr >> 8)];
  const int gj = gd[static_cast<size_t>(e) * kGdStride + (pr & 0xff)];
  unsigned long long key = (1ull << (2 * bits)) - 1ull;  // sentinel sorts last
  if (gi >= 0 && gj >= 0) {
    const unsigned long long r = static_cast<unsigned long long>(gi > gj ? gi : gj);
    const unsigned long long c = static_cast<unsigned long long>(gi > gj ? gj : gi);
    key = (r << bits) | c;
  }
  keys[idx] = key;
}

__global__ void k_mark_unique(long long m, const unsigned long long* __restrict__ keys,
                              unsigned long long sentinel, int* __restrict__ head) {
  const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (i >= m) return;
  const unsigned long long k = keys[i];
  head[i] = (k != sentinel && (i == 0 || keys[i - 1] != k)) ? 1 : 0;
}

__global__ void k_compact_keys(long long m, const unsigned long long* __restrict__ keys,
                               const int* __restrict__ head, const int* __restrict__ pos,
                               int bits, int* __restrict__ row_of, int* __restrict__ col,
                               long long* __restrict__ seg = nullptr) {
  const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (i >= m || !head[i]) return;
  const unsigned long long k = keys[i];
  const int p = pos[i];
  row_of[p] = static_cast<int>(k >> bits);
  col[p] = static_cast<int>(k & ((1ull << bits) - 1ull));
  if (seg) seg[p] = i;
}

// seg[cnt] = number of keys before the first sentinel (sorted keys).
__global__ void k_seg_end(long long m, const unsigned long long* __restrict__ keys,
                          unsigned long long sentinel, int cnt, long long* __restrict__ seg) {
  long long lo = 0, hi = m;  // first index with keys[i] == sentinel
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if (keys[mid] < sentinel) lo = mid + 1;
    else hi = mid;
  }
  seg[cnt] = lo;
}

// Slot sums of a mesh with several element kinds: the run of slot p holds
// the kind groups one after the other, so its entries are visited in
// increasing element index (eid of each pair position; an element
// contributes at most once to a slot). Same skip and mark rules.
__global__ void k_slot_sums_by_element(int nslots, const long long* __restrict__ seg,
                                       const unsigned* __restrict__ src,
                                       const int* __restrict__ eid,
                                       const double* __restrict__ ke, double* __restrict__ val,
                                       int* __restrict__ touched) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nslots) return;
  const long long j0 = seg[p], j1 = seg[p + 1];
  double acc = 0.0;
  int any = 0, last = -1;
  for (long long c = j0; c < j1; ++c) {
    int best = 0x7fffffff;
    long long bj = j0;
    for (long long j = j0; j < j1; ++j) {
      const int ej = eid[src[j]];
      if (ej > last && ej < best) {
        best = ej;
        bj = j;
      }
    }
    if (best == 0x7fffffff) break;  // every element visited
    last = best;
    const double v = ke[src[bj]];
    if (v != 0.0) {
      acc += v;
      any = 1;
    }
  }
  val[p] = acc;
  touched[p] = any;
}

// Slot sums in element order: slot p's contributions are the sorted pair
// positions seg[p] .. seg[p+1]-1, whose entries the element kernels stored at
// ke[src[j]]. Exact zeros are skipped; touched[p] = 1 if any entry was nonzero.
__global__ void k_slot_sums(int nslots, const long long* __restrict__ seg,
                            const unsigned* __restrict__ src, const double* __restrict__ ke,
                            double* __restrict__ val, int* __restrict__ touched) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nslots) return;
  double acc = 0.0;
  int any = 0;
  for (long long j = seg[p]; j < seg[p + 1]; ++j) {
    const double v = ke[src[j]];
    if (v != 0.0) {
      acc += v;
      any = 1;
    }
  }
  val[p] = acc;
  touched[p] = any;
}

// ---- 2'. node-level pattern ------------------------------------------------------------
// The element pairs of pass 2 repeat each node pair's D x D DOF block, so a
// mesh whose nodes all carry D DOFs (solids 3, shells 6) is sorted by node
// pair instead (36 keys per CHEXA8 instead of 300) and each unique node pair
// is expanded into its DOF slots in CSR order. Requires every node's active
// DOFs d < D to have consecutive equation numbers in d order (the DofMap
// numbers node by node), checked by k_node_dofs.

// base[a] = first active equation of node a among DOFs d < D, nact[a] their
// count, dmask[a] bit d set for an active DOF d, node_of_eq[eq] = a; *bad = 1
// if the active equations of a node are not consecutive.
__global__ void k_node_dofs(int nnodes, int D, const int* __restrict__ eq, int* __restrict__ base,
                            int* __restrict__ nact, int* __restrict__ dmask,
                            int* __restrict__ node_of_eq, int* __restrict__ bad) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nnodes) return;
  int b = -1, cnt = 0, prev = -1, mask = 0;
  for (int d = 0; d < D; ++d) {
    const int q = eq[6 * a + d];
    if (q < 0) continue;
    if (b < 0) b = q;
    else if (q != prev + 1) *bad = 1;
    prev = q;
    ++cnt;
    mask |= 1 << d;
    node_of_eq[q] = a;
  }
  base[a] = b;
  nact[a] = cnt;
  dmask[a] = mask;
}

// Node-pair keys of element e at npo[e] .. (its local node pairs a >= b):
// key (larger base << bits) | smaller base, or the sentinel if a node has no
// active DOF; val = ebase[e] << 8 | lr << 4 | lc, lr the local index of the
// row node (larger base), lc of the column node, ebase[e] the position of the
// element's lower entries in ke. Elements are laid out in element order, so
// a stable sort keeps each node pair's contributions in element order.
// *bad = 1 for a repeated node.
__global__ void k_node_pair_keys(int nelem, const int* __restrict__ eptr,
                                 const int* __restrict__ conn, const int* __restrict__ npo,
                                 const int* __restrict__ base, const int* __restrict__ nact,
                                 const unsigned* __restrict__ ebase, int bits,
                                 unsigned long long* __restrict__ keys,
                                 unsigned long long* __restrict__ vals, int* __restrict__ bad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const int o = eptr[e], nn = eptr[e + 1] - o;
  const unsigned long long sentinel = (1ull << (2 * bits)) - 1ull;
  const unsigned long long kb = static_cast<unsigned long long>(ebase[e]) << 8;
  int p = npo[e];
  for (int a = 0; a < nn; ++a) {
    const int A = conn[o + a];
    for (int b = 0; b <= a; ++b, ++p) {
      const int B = conn[o + b];
      unsigned long long key = sentinel;
      int lr = a, lc = b;
      if (nact[A] > 0 && nact[B] > 0) {
        if (a != b && A == B) *bad = 1;
        const unsigned long long ba = static_cast<unsigned long long>(base[A]);
        const unsigned long long bb = static_cast<unsigned long long>(base[B]);
        if (bb > ba) {
          lr = b;
          lc = a;
        }
        key = ((ba > bb ? ba : bb) << bits) | (ba > bb ? bb : ba);
      }
      keys[p] = key;
      vals[p] = kb | static_cast<unsigned long long>((lr << 4) | lc);
    }
  }
}

// Per unique node pair u (sorted by row node, the diagonal pair last):
// cnt[u] = DOFs of its column node unless diagonal; first_of[A] = A's first
// pair, diag_of[A] = its diagonal pair; info[u] = DOF mask of the row node |
// mask of the column node << 8 | (diagonal) << 16.
__global__ void k_np_info(int U, const int* __restrict__ rb, const int* __restrict__ cb,
                          const int* __restrict__ node_of_eq, const int* __restrict__ nact,
                          const int* __restrict__ dmask, int* __restrict__ cnt,
                          int* __restrict__ first_of, int* __restrict__ diag_of,
                          int* __restrict__ info) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= U) return;
  const int A = node_of_eq[rb[u]], B = node_of_eq[cb[u]];
  cnt[u] = (A != B) ? nact[B] : 0;
  if (u == 0 || rb[u - 1] != rb[u]) first_of[A] = u;
  if (A == B) diag_of[A] = u;
  info[u] = dmask[A] | (dmask[B] << 8) | ((A == B) ? (1 << 16) : 0);
}

// off[u]: the column of node pair u's first slot within its rows (the DOFs
// of the row node's earlier column nodes).
__global__ void k_np_off(int U, const int* __restrict__ rb, const int* __restrict__ node_of_eq,
                         const int* __restrict__ pre, const int* __restrict__ first_of,
                         int* __restrict__ off) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= U) return;
  off[u] = pre[u] - pre[first_of[node_of_eq[rb[u]]]];
}

// Lower CSR row lengths: row r of node A (active index i) holds the DOFs of
// A's off-diagonal column nodes and then the diagonal block's first i + 1.
__global__ void k_np_row_len(int n, const int* __restrict__ node_of_eq,
                             const int* __restrict__ base, const int* __restrict__ pre,
                             const int* __restrict__ first_of, const int* __restrict__ diag_of,
                             int* __restrict__ len) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int A = node_of_eq[r];
  if (A < 0 || diag_of[A] < 0) {
    len[r] = 0;
    return;
  }
  len[r] = pre[diag_of[A]] - pre[first_of[A]] + (r - base[A]) + 1;
}

// Slot (i, j) of node pair u: row r = rb + i, column c = cb + j (j <= i on
// the diagonal pair) at rowptr[r] + off[u] + j. Thread t: j = t mod D, then
// u, then i, so a wave's lanes read adjacent entries of the same element
// matrices and write adjacent slots of the same rows. Each slot sums its
// contributions in element order (the sorted run seg[u] .. seg[u + 1]): the
// element's entry of local DOFs (D lr + d_i, D lc + d_j) at ke[kb + t], with
// the skip and mark rules of k_slot_sums. Runs of up to kRunCache elements
// (every node pair of hexa and shell meshes) are loaded at once; longer runs
// one entry at a time.
constexpr int kRunCache = 8;

template <int D>
__device__ __forceinline__ double np_entry(const double* __restrict__ ke, unsigned long long v,
                                           int di, int dj) {
  const int lr = static_cast<int>((v >> 4) & 15u), lc = static_cast<int>(v & 15u);
  const int li = D * lr + di, lj = D * lc + dj;
  const int hi = li > lj ? li : lj, lo = li > lj ? lj : li;
  return ke[(v >> 8) + static_cast<unsigned long long>(hi * (hi + 1) / 2 + lo)];
}

// the k-th set bit of m (k < popcount(m)), or -1
__device__ __forceinline__ int nth_bit(int m, int k) {
  for (int d = 0; d < 8; ++d)
    if ((m >> d) & 1) {
      if (k == 0) return d;
      --k;
    }
  return -1;
}

template <int D>
__global__ void k_np_slot_sums(int U, const int* __restrict__ rb, const int* __restrict__ cb,
                               const long long* __restrict__ seg,
                               const unsigned long long* __restrict__ vals,
                               const int* __restrict__ info, const int* __restrict__ off,
                               const int* __restrict__ rowptr, const double* __restrict__ ke,
                               int* __restrict__ t_row, int* __restrict__ t_col,
                               double* __restrict__ t_val, int* __restrict__ touched) {
  const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<long long>(U) * D * D) return;
  const int j = static_cast<int>(t % D);
  const long long rest = t / D;
  const int u = static_cast<int>(rest % U), i = static_cast<int>(rest / U);
  const int inf = info[u];
  const int mA = inf & 0xff, mB = (inf >> 8) & 0xff;
  const int di = nth_bit(mA, i);
  if (di < 0) return;
  const int jn = ((inf >> 16) & 1) ? i + 1 : __popc(mB);
  if (j >= jn) return;
  const int dj = nth_bit(mB, j);
  const int r = rb[u] + i;
  const long long j0 = seg[u], j1 = seg[u + 1];
  const int L = static_cast<int>(j1 - j0);
  double acc = 0.0;
  int any = 0;
  if (L <= kRunCache) {
    // entries past L repeat the last one (a valid address); only the first L
    // are summed
    double x[kRunCache];
#pragma unroll
    for (int q = 0; q < kRunCache; ++q) x[q] = np_entry<D>(ke, vals[j0 + (q < L ? q : L - 1)], di, dj);
#pragma unroll
    for (int q = 0; q < kRunCache; ++q) {
      if (q < L && x[q] != 0.0) {
        acc += x[q];
        any = 1;
      }
    }
  } else {
    for (long long q = j0; q < j1; ++q) {
      const double x = np_entry<D>(ke, vals[q], di, dj);
      if (x != 0.0) {
        acc += x;
        any = 1;
      }
    }
  }
  const int p = rowptr[r] + off[u] + j;
  t_row[p] = r;
  t_col[p] = cb[u] + j;
  t_val[p] = acc;
  touched[p] = any;
}

// row_ptr from a row-sorted entry list (handles empty rows).
__global__ void k_row_ptr(int nnz, int n, const int* __restrict__ row_of, int* __restrict__ rp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nnz) return;
  const int prev = (i == 0) ? -1 : row_of[i - 1];
  const int cur = (i == nnz) ? n : row_of[i];
  for (int r = prev + 1; r <= cur; ++r) rp[r] = i;
}

// ---- 3. element matrices -------------------------------------------------------------------
struct M3d {
  double a[3][3];
};

__device__ __forceinline__ double det3(const M3d& m) {
  return m.a[0][0] * (m.a[1][1] * m.a[2][2] - m.a[1][2] * m.a[2][1]) -
         m.a[1][0] * (m.a[0][1] * m.a[2][2] - m.a[0][2] * m.a[2][1]) +
         m.a[2][0] * (m.a[0][1] * m.a[1][2] - m.a[0][2] * m.a[1][1]);
}

// Inverse by cofactors with its own determinant from the first-row
// expansion, in the operation order of the CPU oracle's inv3 (bit-identical
// with contraction off). `bad` replaces a non-positive determinant by 1.
__device__ __forceinline__ M3d inv3(const M3d& m, bool bad) {
  const double c00 = m.a[1][1] * m.a[2][2] - m.a[1][2] * m.a[2][1];
  const double c10 = m.a[1][2] * m.a[2][0] - m.a[1][0] * m.a[2][2];
  const double c20 = m.a[1][0] * m.a[2][1] - m.a[1][1] * m.a[2][0];
  const double det = m.a[0][0] * c00 + m.a[0][1] * c10 + m.a[0][2] * c20;
  const double id = 1.0 / (bad ? 1.0 : det);
  M3d r;
  r.a[0][0] = c00 * id;
  r.a[1][0] = c10 * id;
  r.a[2][0] = c20 * id;
  r.a[0][1] = (m.a[0][2] * m.a[2][1] - m.a[0][1] * m.a[2][2]) * id;
  r.a[1][1] = (m.a[0][0] * m.a[2][2] - m.a[0][2] * m.a[2][0]) * id;
  r.a[2][1] = (m.a[0][1] * m.a[2][0] - m.a[0][0] * m.a[2][1]) * id;
  r.a[0][2] = (m.a[0][1] * m.a[1][2] - m.a[0][2] * m.a[1][1]) * id;
  r.a[1][2] = (m.a[0][2] * m.a[1][0] - m.a[0][0] * m.a[1][2]) * id;
  r.a[2][2] = (m.a[0][0] * m.a[1][1] - m.a[0][1] * m.a[1][0]) * id;
  return r;
}

// Parametric shape derivatives (reference CHexa8::shape_functions
// solid_elements.cpp:76-93 and CPenta6::shape_functions 510-548).
template <int K>
__device__ __forceinline__ void shape_d(double r0, double r1, double r2, double (&d)[Kind<K>::NN][3]) {
  if constexpr (Kind<K>::NN == 8) {
    constexpr double xn[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    constexpr double en[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    constexpr double zn[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      d[i][0] = 0.125 * xn[i] * (1 + en[i] * r1) * (1 + zn[i] * r2);
      d[i][1] = 0.125 * en[i] * (1 + xn[i] * r0) * (1 + zn[i] * r2);
      d[i][2] = 0.125 * zn[i] * (1 + xn[i] * r0) * (1 + en[i] * r1);
    }
  } else if constexpr (Kind<K>::NN == 6) {
    const double L3 = 1.0 - r0 - r1, zm = (1.0 - r2) * 0.5, zp = (1.0 + r2) * 0.5;
    const double a[6] = {zm, 0.0, -zm, zp, 0.0, -zp};
    const double b[6] = {0.0, zm, -zm, 0.0, zp, -zp};
    const double c[6] = {-r0 * 0.5, -r1 * 0.5, -L3 * 0.5, r0 * 0.5, r1 * 0.5, L3 * 0.5};
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      d[i][0] = a[i];
      d[i][1] = b[i];
      d[i][2] = c[i];
    }
  }
}

template <int NN>
__device__ __forceinline__ M3d jac(const double (&X)[NN][3], const double (&d)[NN][3]) {
  M3d J;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) J.a[i][j] = 0.0;
#pragma unroll
  for (int n = 0; n < NN; ++n)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) J.a[i][j] += d[n][i] * X[n][j];
  return J;
}

// An isotropic material matrix in strain order (xx, yy, zz, xy, yz, zx): its
// normal 3x3 block (diagonal nd, off-diagonal no) and the shear diagonal s.
struct Dmat {
  double nd, no, s;
};

// One entry of B^T D B, B(:, i) the strain column of displacement component p
// with gradient a, B(:, j) component q with gradient b (build_B layout,
// solid_elements.cpp:111-125; an enhanced mode column has the same shape,
// fill_G_col 1140-1151), in the order of the reference's product
// (sum_k B[k][i] (sum_l D[k][l] B[l][j])), every sum from 0.0 in increasing
// index: each (D B)[k][j] has a single nonzero product, and the row sum
// visits the normal row p, then the two shear rows of p in increasing order.
// Bit-identical to the CPU oracle's add_BtDB with contraction off (terms that
// are exact zeros there leave the sums unchanged).
__device__ __forceinline__ double btdb6(const Dmat& D, const double* a, int p, const double* b,
                                        int q) {
  const double s3 = (q == 0) ? D.s * b[1] : (q == 1 ? D.s * b[0] : 0.0);
  const double s4 = (q == 1) ? D.s * b[2] : (q == 2 ? D.s * b[1] : 0.0);
  const double s5 = (q == 0) ? D.s * b[2] : (q == 2 ? D.s * b[0] : 0.0);
  double v = a[p] * ((p == q ? D.nd : D.no) * b[q]);
  if (p == 0) {
    v += a[1] * s3;
    v += a[2] * s5;
  } else if (p == 1) {
    v += a[0] * s3;
    v += a[2] * s4;
  } else {
    v += a[1] * s4;
    v += a[0] * s5;
  }
  return v;
}

__device__ __forceinline__ const int* g_of(const int* gd, int e) {
  return gd + static_cast<size_t>(e) * kGdStride;
}

// det3 with the operation order of the oracle (expansion along column 0).
__device__ double det3_nc(double m00, double m01, double m02, double m10, double m11, double m12,
                          double m20, double m21, double m22) {
#pragma clang fp contract(off)
  const double h0 = m00 * (m11 * m22 - m12 * m21);
  const double h1 = m10 * (m01 * m22 - m02 * m21);
  const double h2 = m20 * (m01 * m12 - m02 * m11);
  return h0 - h1 + h2;
}

template <int SU>
__device__ void tet_entries(const double (&X)[4][3], double E, double nu, int lane, const int* g,
                            double (&kuu)[SU], bool& bad) {
#pragma clang fp contract(off)
  double A[4][4];
  for (int i = 0; i < 4; ++i) {
    A[i][0] = 1.0;
    for (int d = 0; d < 3; ++d) A[i][d + 1] = X[i][d];
  }
  auto minor_det = [&](int sr, int sc) {
    double m[3][3];
    int ri = 0;
    for (int r = 0; r < 4; ++r) {
      if (r == sr) continue;
      int cj = 0;
      for (int c = 0; c < 4; ++c) {
        if (c == sc) continue;
        m[ri][cj++] = A[r][c];
      }
      ++ri;
    }
    return det3_nc(m[0][0], m[0][1], m[0][2], m[1][0], m[1][1], m[1][2], m[2][0], m[2][1], m[2][2]);
  };
  double V6 = 0.0;
  for (int c = 0; c < 4; ++c) V6 += ((c % 2) ? -1.0 : 1.0) * A[0][c] * minor_det(0, c);
  const double V = V6 / 6.0;
  if (!(V > 0)) bad = true;
  double dn[4][3];
  for (int n = 0; n < 4; ++n)
    for (int d = 0; d < 3; ++d) {
      const int i = d + 1;  // cof(i, n) = (-1)^(i+n) minor(n, i)
      dn[n][d] = ((((i + n) % 2) ? -1.0 : 1.0) * minor_det(n, i)) / V6;
    }
  const double lam = E * nu / ((1 + nu) * (1 - 2 * nu));
  const double mu = E / (2 * (1 + nu));
  double D[6][6];
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) D[a][b] = 0.0;
  D[0][0] = lam + 2 * mu; D[0][1] = lam; D[0][2] = lam;
  D[1][0] = lam; D[1][1] = lam + 2 * mu; D[1][2] = lam;
  D[2][0] = lam; D[2][1] = lam; D[2][2] = lam + 2 * mu;
  D[3][3] = mu; D[4][4] = mu; D[5][5] = mu;
  auto bcol = [&](int j, double (&b)[6]) {
    const int n = j / 3, c = j % 3;
    for (int k = 0; k < 6; ++k) b[k] = 0.0;
    const double x = dn[n][0], y = dn[n][1], z = dn[n][2];
    if (c == 0) { b[0] = x; b[3] = y; b[5] = z; }
    else if (c == 1) { b[1] = y; b[3] = x; b[4] = z; }
    else { b[2] = z; b[4] = y; b[5] = x; }
  };
#pragma unroll
  for (int q = 0; q < SU; ++q) {
    const int t = lane + 64 * q;
    if (t < 78) {
      const unsigned short pr = pair_at<12>(t);
      int i = pr >> 8, j = pr & 0xff;
      if (g[i] < g[j]) { const int tmp = i; i = j; j = tmp; }  // row = larger global index
      double bi[6], bj[6];
      bcol(i, bi);
      bcol(j, bj);
      double sum = 0.0;
      for (int k = 0; k < 6; ++k) {
        double db = 0.0;
        for (int l = 0; l < 6; ++l) db += D[k][l] * bj[l];
        sum += bi[k] * db;
      }
      kuu[q] = sum * V;
    }
  }
}

// Per-wave geometry of the current point (solid kernels).
template <int NN>
struct Geo {
  double X[NN][3];    // node coordinates
  double d[NN][3];    // parametric shape derivatives at the point
  double J[9];        // Jacobian, row-major
  double dN[NN][3];   // physical gradients at the point
  double dN0[NN][3];  // physical gradients at the centroid (SRI volumetric part)
  double gk[3][3];    // enhanced-mode gradients at the point (EAS)
  double Ji[9];       // inverse Jacobian at the point, row-major
  double J0i[9];      // inverse centroidal Jacobian
};

// shape_d for one node (the same expressions, so the same values).
template <int K>
__device__ __forceinline__ void shape_d_node(int i, double r0, double r1, double r2, double (&d)[3]) {
  if constexpr (Kind<K>::NN == 8) {
    const double xn = (i == 1 || i == 2 || i == 5 || i == 6) ? 1.0 : -1.0;
    const double en = (i == 2 || i == 3 || i == 6 || i == 7) ? 1.0 : -1.0;
    const double zn = (i >= 4) ? 1.0 : -1.0;
    d[0] = 0.125 * xn * (1 + en * r1) * (1 + zn * r2);
    d[1] = 0.125 * en * (1 + xn * r0) * (1 + zn * r2);
    d[2] = 0.125 * zn * (1 + xn * r0) * (1 + en * r1);
  } else {
    const double L3 = 1.0 - r0 - r1, zm = (1.0 - r2) * 0.5, zp = (1.0 + r2) * 0.5;
    const double zs = (i < 3) ? zm : zp;  // node i's bottom/top factor
    const int k = i % 3;
    d[0] = (k == 0) ? ((i < 3) ? zm : zp) : (k == 1 ? 0.0 : -zs);
    d[1] = (k == 1) ? zs : (k == 0 ? 0.0 : -zs);
    const double Lk = (k == 0) ? r0 : (k == 1 ? r1 : L3);
    d[2] = (i < 3) ? -Lk * 0.5 : Lk * 0.5;
  }
}

template <int K>
struct ElemLds {
  static constexpr int ND = Kind<K>::ND;
  static constexpr int WORDS = Kind<K>::EAS ? (ND * 9 + 81 + 9 * ND) : 1;
};

// MODE 0 scatters K_e into the lower CSR; MODE 1 writes
// out[e] = -u_e^T (dK_e/dE_e) v_e = -u_e^T K_e v_e / E_e (K_e is linear in E);
// MODE 2 adds the element's internal force K_e u_e to out[6 * node + dof]
// (every DOF of the element's nodes, constrained ones included; u_e is 0 at
// DOFs without an equation).
// Resident waves per SIMD asked of the compiler: 4 for assembly (MODE 0; it
// then holds 128 VGPRs and spills the rest to scratch, and the extra waves
// hide more latency than the spills cost: C2 6.5 -> 5.4 ms, 2 waves at 239
// VGPRs without the bound), 2 for the contractions (C2 adjoint 12.7 -> 7.9 ms,
// 1 wave at 257 registers without it).
template <int K, int MODE = 0>
__global__ __launch_bounds__(64 * kElemWaves, MODE == 0 ? 4 : 2) void k_element_stiffness(
    int nsel, const int* __restrict__ sel, const int* __restrict__ eptr,
    const int* __restrict__ conn, const double* __restrict__ xyz, const double* __restrict__ Ev,
    const double* __restrict__ nuv, const int* __restrict__ gd, const int* __restrict__ rp,
    const int* __restrict__ ci, double* __restrict__ val, int* __restrict__ touched,
    int* __restrict__ err, const double* __restrict__ uvec = nullptr,
    const double* __restrict__ vvec = nullptr, double* __restrict__ out = nullptr) {
  using KD = Kind<K>;
  constexpr int NN = KD::NN, ND = KD::ND, NL = KD::NL;
  __shared__ double lds_all[kElemWaves][ElemLds<K>::WORDS];
  __shared__ double facc_all[kElemWaves][MODE == 2 ? KD::ND : 1];
  __shared__ Geo<NN> geo_all[kElemWaves];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* lds = lds_all[wave];
  double* facc = facc_all[wave];
  for (int base = blockIdx.x * kElemWaves; base < nsel; base += gridDim.x * kElemWaves) {
    const int s = base + wave;
    const bool active = s < nsel;
    const int e = active ? sel[s] : sel[base];
    const int o = eptr[e];
    const double E = Ev[e], nu = nuv[e];
    const double lam = E * nu / ((1 + nu) * (1 - 2 * nu));
    const double mu = E / (2 * (1 + nu));
    // isotropic_D (solid_elements.cpp:17-27); SRI splits it into the
    // deviatoric D - Dvol at the Gauss points and Dvol (lam in the normal
    // block) at the centroid (95-180)
    Dmat Dfull, Ddev, Dvol;
    Dfull.nd = lam + 2 * mu;
    Dfull.no = lam;
    Dfull.s = mu;
    Ddev.nd = Dfull.nd - lam;
    Ddev.no = lam - lam;
    Ddev.s = mu - 0.0;
    Dvol.nd = lam;
    Dvol.no = lam;
    Dvol.s = 0.0;
    // lane-owned lower pairs, oriented so the row has the larger global index
    // (add_element_stiffness takes ke(i, j) with gdof[i] >= gdof[j],
    // sparse_matrix.cpp:16-35)
    int ei[KD::SU], ej[KD::SU];
#pragma unroll
    for (int q = 0; q < KD::SU; ++q) {
      const int t = lane + 64 * q;
      const unsigned short pr = pair_at<ND>(t < NL ? t : 0);
      int i = pr >> 8, j = pr & 0xff;
      const int* gg = g_of(gd, e);
      if (gg[i] < gg[j]) {
        const int tmp = i;
        i = j;
        j = tmp;
      }
      ei[q] = i;
      ej[q] = j;
    }
    double kuu[KD::SU];
#pragma unroll
    for (int q = 0; q < KD::SU; ++q) kuu[q] = 0.0;
    double kua[KD::SA];
#pragma unroll
    for (int q = 0; q < KD::SA; ++q) kua[q] = 0.0;
    double kaa = 0.0;
    bool bad = false;

    if constexpr (K == kTet4) {
      // CST tetrahedron (solid_elements.cpp:312-384): gradients from the
      // cofactors of A4 = [1 x y z], Ke = V * B^T D B, evaluated in the same
      // operation order as the CPU oracle without FMA contraction, so that
      // entries which cancel to exactly 0.0 there do so here too (the CSR
      // pattern drops exact zeros).
      double X[NN][3];
#pragma unroll
      for (int a = 0; a < NN; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) X[a][d] = xyz[3 * conn[o + a] + d];
      tet_entries<KD::SU>(X, E, nu, lane, g_of(gd, e), kuu, bad);
    } else {
      // Gauss rule: hexa 2x2x2 (weight 1), penta 3 triangle x 2 axial (1/6).
      // The geometry of each point is formed once per wave in LDS (shape
      // derivatives by node lanes, J by entry lanes, physical gradients by
      // DOF lanes); every lane then reads the gradients of its entries, so
      // no lane holds whole per-node arrays (which spilled to scratch).
      constexpr int NG = (NN == 8) ? 8 : 6;
      const double gp = 0.57735026918962584;  // 1/sqrt(3)
      Geo<NN>& G = geo_all[wave];
      // the shape derivatives of node `lane` at (r0, r1, r2) into G.d, then J
      // (sum over nodes in node order, the oracle's jacobian) into G.J;
      // returns the point's inverse Jacobian and determinant in all lanes
      auto point_geometry = [&](double r0, double r1, double r2, M3d& Ji, double& detJ) {
        if (lane < NN) {
          double dl[3];
          shape_d_node<K>(lane, r0, r1, r2, dl);
          G.d[lane][0] = dl[0];
          G.d[lane][1] = dl[1];
          G.d[lane][2] = dl[2];
        }
        __syncthreads();
        if (lane < 9) {
          const int ia = lane / 3, ib = lane % 3;
          double sj = 0.0;
          for (int n = 0; n < NN; ++n) sj += G.d[n][ia] * G.X[n][ib];
          G.J[lane] = sj;
        }
        __syncthreads();
        M3d J;
#pragma unroll
        for (int q = 0; q < 9; ++q) J.a[q / 3][q % 3] = G.J[q];
        detJ = det3(J);
        if (!(detJ > 0)) bad = true;
        Ji = inv3(J, bad);
        // lanes index the inverse by their own row: through LDS, not a
        // private array (which would live in scratch memory)
        if (lane == 0)
#pragma unroll
          for (int q = 0; q < 9; ++q) G.Ji[q] = Ji.a[q / 3][q % 3];
        __syncthreads();
      };
      // physical gradients dN[n][c] (the oracle's phys_grad) by DOF lanes
      auto grads = [&]() {
        if (lane < ND) {
          const int n = lane / 3, c = lane % 3;
          G.dN[n][c] = G.Ji[3 * c] * G.d[n][0] + G.Ji[3 * c + 1] * G.d[n][1] +
                       G.Ji[3 * c + 2] * G.d[n][2];
        }
      };
      if (lane < ND) G.X[lane / 3][lane % 3] = xyz[3 * conn[o + lane / 3] + lane % 3];
      __syncthreads();
      // centroidal Jacobian (EAS mapping / SRI volumetric point)
      M3d J0i;
      double detJ0;
      if constexpr (NN == 8) point_geometry(0.0, 0.0, 0.0, J0i, detJ0);
      else point_geometry(1.0 / 3.0, 1.0 / 3.0, 0.0, J0i, detJ0);
      if (lane < 9) G.J0i[lane] = G.Ji[lane];
      if constexpr (K == kHexSri) {
        // volumetric part at the centroid, weight 8 (solid_elements.cpp:153-177),
        // added after the deviatoric Gauss points below
        grads();
        __syncthreads();
        if (lane < ND) G.dN0[lane / 3][lane % 3] = G.dN[lane / 3][lane % 3];
      }
      __syncthreads();
#pragma unroll 1
      for (int g = 0; g < NG; ++g) {
        double r0, r1, r2, wq;
        if constexpr (NN == 8) {
          r0 = (g & 4) ? gp : -gp;
          r1 = (g & 2) ? gp : -gp;
          r2 = (g & 1) ? gp : -gp;
          wq = 1.0;
        } else {
          const int ti = g >> 1;
          r0 = (ti == 0) ? 2.0 / 3.0 : 1.0 / 6.0;
          r1 = (ti == 1) ? 2.0 / 3.0 : 1.0 / 6.0;
          r2 = (g & 1) ? gp : -gp;
          wq = 1.0 / 6.0;
        }
        M3d Ji;
        double detJ;
        point_geometry(r0, r1, r2, Ji, detJ);
        grads();
        if constexpr (KD::EAS) {
          // enhanced-mode gradients g_k (scaled by detJ0/detJ):
          //   hexa: rows of J0^{-1} times (xi, eta, zeta)   (solid_elements.cpp:1131-1133)
          //   penta: columns of J0^{-1} times (L1-1/3, L2-1/3, zeta) (1399-1410)
          if (lane >= 32 && lane < 41) {
            const double sc = detJ0 / detJ;
            const int k = (lane - 32) / 3, c = (lane - 32) % 3;
            double gv;
            if constexpr (NN == 8) {
              const double prm = k == 0 ? r0 : (k == 1 ? r1 : r2);
              gv = G.J0i[3 * k + c] * prm * sc;
            } else {
              const double f = k == 0 ? r0 - 1.0 / 3.0 : (k == 1 ? r1 - 1.0 / 3.0 : r2);
              gv = G.J0i[3 * c + k] * f * sc;
            }
            G.gk[k][c] = gv;
          }
        }
        __syncthreads();
        const double w = detJ * wq;
        const Dmat& Dg = (K == kHexSri) ? Ddev : Dfull;  // SRI: deviatoric part here
#pragma unroll
        for (int q = 0; q < KD::SU; ++q) {
          const int t = lane + 64 * q;
          if (t < NL) {
            const int i = ei[q], j = ej[q];
            kuu[q] += btdb6(Dg, G.dN[i / 3], i % 3, G.dN[j / 3], j % 3) * w;
          }
        }
        if constexpr (KD::EAS) {
#pragma unroll
          for (int q = 0; q < KD::SA; ++q) {
            const int t = lane + 64 * q;
            if (t < ND * 9) {
              const int i = t / 9, m = t % 9;
              kua[q] += btdb6(Dfull, G.dN[i / 3], i % 3, G.gk[m / 3], m % 3) * w;
            }
          }
          if (lane < 45) {
            const unsigned short pr = pair_at<9>(lane);
            const int m1 = pr >> 8, m2 = pr & 0xff;
            kaa += btdb6(Dfull, G.gk[m1 / 3], m1 % 3, G.gk[m2 / 3], m2 % 3) * w;
          }
        }
        __syncthreads();
      }
      if constexpr (K == kHexSri) {
        const double w = detJ0 * 8.0;
#pragma unroll
        for (int q = 0; q < KD::SU; ++q) {
          const int t = lane + 64 * q;
          if (t < NL) {
            const int i = ei[q], j = ej[q];
            kuu[q] += btdb6(Dvol, G.dN0[i / 3], i % 3, G.dN0[j / 3], j % 3) * w;
          }
        }
      }
    }

    if constexpr (KD::EAS) {
      // static condensation Ke = Kuu - Kua Kaa^{-1} Kua^T, Kaa factored in LDS
      double* sKua = lds;               // [ND][9]
      double* sKaa = lds + ND * 9;      // [9][9] -> L in place
      double* sX = sKaa + 81;           // [9][ND]
#pragma unroll
      for (int q = 0; q < KD::SA; ++q) {
        const int t = lane + 64 * q;
        if (t < ND * 9) sKua[t] = kua[q];
      }
      if (lane < 45) {
        const unsigned short pr = pair_at<9>(lane);
        const int m1 = pr >> 8, m2 = pr & 0xff;
        sKaa[m1 * 9 + m2] = kaa;
        sKaa[m2 * 9 + m1] = kaa;
      }
      __syncthreads();
      if (lane == 0) {
        for (int j = 0; j < 9; ++j) {
          double dj = sKaa[j * 9 + j];
          for (int k = 0; k < j; ++k) dj -= sKaa[j * 9 + k] * sKaa[j * 9 + k];
          if (!(dj > 0.0)) {
            dj = 1.0;
            if (active) atomicOr(&err[0], kErrKaa);
          }
          dj = sqrt(dj);
          sKaa[j * 9 + j] = dj;
          for (int i = j + 1; i < 9; ++i) {
            double sv = sKaa[i * 9 + j];
            for (int k = 0; k < j; ++k) sv -= sKaa[i * 9 + k] * sKaa[j * 9 + k];
            sKaa[i * 9 + j] = sv / dj;
          }
        }
      }
      __syncthreads();
      if (lane < ND) {
        double y[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
          double sv = sKua[lane * 9 + i];
          for (int k = 0; k < i; ++k) sv -= sKaa[i * 9 + k] * y[k];
          y[i] = sv / sKaa[i * 9 + i];
        }
#pragma unroll
        for (int i = 8; i >= 0; --i) {
          double sv = y[i];
          for (int k = i + 1; k < 9; ++k) sv -= sKaa[k * 9 + i] * y[k];
          y[i] = sv / sKaa[i * 9 + i];
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) sX[i * ND + lane] = y[i];
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < KD::SU; ++q) {
        const int t = lane + 64 * q;
        if (t < NL) {
          const int i = ei[q], j = ej[q];
          double sv = 0.0;
#pragma unroll
          for (int k = 0; k < 9; ++k) sv += sKua[i * 9 + k] * sX[k * ND + j];
          kuu[q] -= sv;
        }
      }
      __syncthreads();
    }

    if constexpr (MODE == 1) {
      const int* g = gd + static_cast<size_t>(e) * kGdStride;
      double part = 0.0;
#pragma unroll
      for (int q = 0; q < KD::SU; ++q) {
        const int t = lane + 64 * q;
        if (t < NL) {
          const unsigned short pr = pair_at<ND>(t);
          const int i = pr >> 8, j = pr & 0xff;
          const int gi = g[i], gj = g[j];
          if (gi >= 0 && gj >= 0) {
            double w = uvec[gi] * vvec[gj];
            if (i != j) w += uvec[gj] * vvec[gi];
            part += kuu[q] * w;
          }
        }
      }
      for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off, 64);
      if (active && lane == 0) {
        out[e] = -part / E;
        if (bad) {
          atomicOr(&err[0], kErrJacobian);
          atomicMin(&err[1], e);
        }
      }
    }
    if constexpr (MODE == 2) {
      // f_e = K_e u_e from the lane-owned lower entries, summed in LDS
      const int* g = gd + static_cast<size_t>(e) * kGdStride;
      if (lane < ND) facc[lane] = 0.0;
      __syncthreads();
#pragma unroll
      for (int q = 0; q < KD::SU; ++q) {
        const int t = lane + 64 * q;
        if (t < NL) {
          const unsigned short pr = pair_at<ND>(t);
          const int i = pr >> 8, j = pr & 0xff;
          const double ui = g[i] >= 0 ? uvec[g[i]] : 0.0, uj = g[j] >= 0 ? uvec[g[j]] : 0.0;
          atomicAdd(&facc[i], kuu[q] * uj);
          if (i != j) atomicAdd(&facc[j], kuu[q] * ui);
        }
      }
      __syncthreads();
      if (active && lane < ND) atomicAdd(&out[6 * conn[o + lane / 3] + lane % 3], facc[lane]);
      if (active && lane == 0 && bad) {
        atomicOr(&err[0], kErrJacobian);
        atomicMin(&err[1], e);
      }
      __syncthreads();
    }
    if (MODE == 0 && active) {
      if (bad && lane == 0) {
        atomicOr(&err[0], kErrJacobian);
        atomicMin(&err[1], e);
      }
      // entry t of element s at val[s * NL + t] (its pair position; the
      // slot sums read only pairs whose DOFs are both free)
      double* dst = val + static_cast<size_t>(s) * NL;
#pragma unroll
      for (int q = 0; q < KD::SU; ++q) {
        const int t = lane + 64 * q;
        if (t < NL) dst[t] = kuu[q];
      }
    }
  }
}

// ---- 4. compaction -------------------------------------------------------------------------------
__global__ void k_compact_csr(int nnz, const int* __restrict__ touched, const int* __restrict__ pos,
                              const int* __restrict__ ci, const double* __restrict__ v,
                              const int* __restrict__ row_of, int* __restrict__ ci2,
                              double* __restrict__ v2, int* __restrict__ row2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz || !touched[i]) return;
  const int p = pos[i];
  ci2[p] = ci[i];
  v2[p] = v[i];
  row2[p] = row_of[i];
}

// Jacobi diagonal from the lower CSR (diagonal is the last entry of each row);
// err bit 8 zero diagonal, bit 16 missing diagonal, err[2] = first bad row.
__global__ void k_diag_from_lower(int n, int n_pad, const int* __restrict__ rp,
                                  const int* __restrict__ ci, const double* __restrict__ v,
                                  double* __restrict__ d64, float* __restrict__ d32,
                                  int* __restrict__ err) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_pad) return;
  double d = 1.0;
  if (r < n) {
    const int e = rp[r + 1] - 1;
    if (e < rp[r] || ci[e] != r) {
      atomicOr(&err[0], 16);
      atomicMin(&err[2], r);
    } else {
      d = v[e];
      if (d == 0.0) {
        atomicOr(&err[0], 8);
        atomicMin(&err[2], r);
      }
    }
  }
  d64[r] = d;
  d32[r] = static_cast<float>(d);
}

// Row checks of a caller's lower CSR, made in the order of the host loop of
// build_jacobi_diagonal (cuda_pcg_backend_impl.hpp:2076-2101) with the
// argument checks of vbh_set_csr added: the first failing row decides,
// and in a row its first failing entry, the missing-diagonal check coming
// after the row's last entry. first_err = min over rows of row << 3 | code,
// code 1 column out of range, 2 entry above the diagonal, 3 zero diagonal,
// 4 missing diagonal. Also writes each entry's row and sets *unsorted when a
// row's columns are not strictly increasing (a duplicate included).
__global__ void k_check_lower_rows(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                   const double* __restrict__ v, int* __restrict__ row_of,
                                   unsigned long long* __restrict__ first_err,
                                   int* __restrict__ unsorted) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int k0 = rp[r], k1 = rp[r + 1];
  int code = 0, prev = -1;
  bool found = false, sorted = true;
  for (int k = k0; k < k1; ++k) {
    row_of[k] = r;
    const int c = ci[k];
    if (code) continue;
    if (c < 0 || c >= n) {
      code = 1;
    } else if (c > r) {
      code = 2;
    } else {
      sorted = sorted && c > prev;
      prev = c;
      if (c == r && !found) {
        if (v[k] == 0.0) code = 3;
        found = true;
      }
    }
  }
  if (!code && !found) code = 4;
  if (code) atomicMin(first_err, (static_cast<unsigned long long>(r) << 3) | static_cast<unsigned>(code));
  if (!sorted) atomicOr(unsorted, 1);
}

__global__ void k_init_check(unsigned long long* first_err, int* unsorted) {
  *first_err = ~0ull;
  *unsorted = 0;
}

// (slice, block-column offset) keys of the relative sliced ELL: block row I
// covers rows B*I .. B*I+B-1 and so lies in one slice or two; each of its
// blocks q gives the key slice << bits | (bci[q] - I + nbr) for each of them
// (the sentinel where the row lies in one slice only).
__global__ void k_rel_keys(int nbr, int B, const int* __restrict__ brp, const int* __restrict__ bci,
                           int bits, unsigned long long sentinel,
                           unsigned long long* __restrict__ keys) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nbr) return;
  const unsigned long long s0 = static_cast<unsigned>((B * I) / kSliceRows);
  const unsigned long long s1 = static_cast<unsigned>((B * I + B - 1) / kSliceRows);
  for (int q = brp[I]; q < brp[I + 1]; ++q) {
    const unsigned long long off = static_cast<unsigned>(bci[q] - I + nbr);
    keys[2 * static_cast<size_t>(q)] = (s0 << bits) | off;
    keys[2 * static_cast<size_t>(q) + 1] = s1 != s0 ? ((s1 << bits) | off) : sentinel;
  }
}

// 6x6 block keys from a 3x3 block pattern (both triangles): 6x6 block
// (I/2, J/2) holds 3x3 block (I, J).
__global__ void k_coarse_keys(int nnzb, const int* __restrict__ brow, const int* __restrict__ bcol,
                              int bits, unsigned long long* __restrict__ keys) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nnzb) return;
  keys[q] = (static_cast<unsigned long long>(brow[q] >> 1) << bits) |
            static_cast<unsigned long long>(bcol[q] >> 1);
}

// ---- BSR (full symmetric) from the lower CSR ----------------------------------------------------
__global__ void k_block_keys(int nnz, int B, const int* __restrict__ row_of,
                             const int* __restrict__ ci, unsigned long long* __restrict__ keys,
                             int bits) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  const unsigned long long I = static_cast<unsigned long long>(row_of[i] / B);
  const unsigned long long J = static_cast<unsigned long long>(ci[i] / B);
  keys[2 * static_cast<size_t>(i)] = (I << bits) | J;
  keys[2 * static_cast<size_t>(i) + 1] = (J << bits) | I;
}

template <int B>
__global__ void k_bsr_fill(int nnz, int n, const int* __restrict__ row_of, const int* __restrict__ ci,
                           const double* __restrict__ v, const int* __restrict__ brp,
                           const int* __restrict__ bci, double* __restrict__ bv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  const int r = row_of[i], c = ci[i];
  const double x = v[i];
  {
    const int I = r / B, J = c / B;
    const int o = brp[I], nb = brp[I + 1] - o;
    const int slot = find_slot(brp, bci, I, J);
    if (slot >= 0)
      bv[static_cast<size_t>(B * B) * o + static_cast<size_t>((r % B) * B + (c % B)) * nb + (slot - o)] = x;
  }
  if (r != c) {
    const int I = c / B, J = r / B;
    const int o = brp[I], nb = brp[I + 1] - o;
    const int slot = find_slot(brp, bci, I, J);
    if (slot >= 0)
      bv[static_cast<size_t>(B * B) * o + static_cast<size_t>((c % B) * B + (r % B)) * nb + (slot - o)] = x;
  }
}

template <int B>
__global__ void k_bsr_pad_diag(int n, int nbr, const int* __restrict__ brp,
                               const int* __restrict__ bci, double* __restrict__ bv) {
  const int r = n + blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nbr * B) return;
  const int I = r / B;
  const int o = brp[I], nb = brp[I + 1] - o;
  const int slot = find_slot(brp, bci, I, I);
  if (slot >= 0)
    bv[static_cast<size_t>(B * B) * o + static_cast<size_t>((r % B) * B + (r % B)) * nb + (slot - o)] = 1.0;
}

__global__ void k_cast(long long m, const double* __restrict__ a, float* __restrict__ b) {
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < m;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    b[i] = static_cast<float>(a[i]);
}

__global__ void k_zero_i32(long long m, int* __restrict__ a) {
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < m;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    a[i] = 0;
}

__global__ void k_set_i32(long long m, int* __restrict__ a, int v) {
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < m;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    a[i] = v;
}

__global__ void k_zero_f64(long long m, double* __restrict__ a) {
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < m;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    a[i] = 0.0;
}

__global__ void k_init_err(int* err) {
  err[0] = 0;
  err[1] = 0x7fffffff;
  err[2] = 0x7fffffff;
  err[3] = 0;
}

// ---- DOF numbering -------------------------------------------------------------------------------
__global__ void k_dof_active(int nelem, const int* __restrict__ etype, const int* __restrict__ eptr,
                             const int* __restrict__ conn, int* __restrict__ mask) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const int bits = (etype[e] == VBH_CQUAD4) ? 0x3f : 0x07;  // mark_element_dofs 39-68
  for (int k = eptr[e]; k < eptr[e + 1]; ++k) atomicOr(&mask[conn[k]], bits);
}

__global__ void k_dof_spc(int nspc, const int* __restrict__ sn, const int* __restrict__ sm,
                          int* __restrict__ spc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nspc) return;
  atomicOr(&spc[sn[i]], sm[i] & 0x3f);
}

__global__ void k_dof_count(int nn, const int* __restrict__ order, const int* __restrict__ mask,
                            const int* __restrict__ spc, int* __restrict__ cnt) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nn) return;
  const int i = order[s];
  cnt[s] = __popc(mask[i] & ~spc[i]);
}

__global__ void k_dof_assign(int nn, const int* __restrict__ order, const int* __restrict__ mask,
                             const int* __restrict__ spc, const int* __restrict__ base,
                             int* __restrict__ eq) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nn) return;
  const int i = order[s];
  const int f = mask[i] & ~spc[i];
  int k = base[s];
  for (int d = 0; d < 6; ++d) eq[6 * i + d] = ((f >> d) & 1) ? k++ : -1;
}

int bits_for(long long n) {
  int b = 1;
  while ((1ll << b) <= n) ++b;
  return b;
}

template <typename T>
void dzero(DBuf<T>& b, hipStream_t s);
template <>
void dzero<int>(DBuf<int>& b, hipStream_t s) {
  if (b.n) hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(b.n, 1024)), dim3(256), 0, s,
                              static_cast<long long>(b.n), b.p);
}
template <>
void dzero<double>(DBuf<double>& b, hipStream_t s) {
  if (b.n) hipLaunchKernelGGL(k_zero_f64, dim3(grid_for(b.n, 1024)), dim3(256), 0, s,
                              static_cast<long long>(b.n), b.p);
}

// rocPRIM helpers ----------------------------------------------------------------------------------
void sort_keys(DBuf<unsigned long long>& in, DBuf<unsigned long long>& out, size_t m, int end_bit,
               hipStream_t s) {
  size_t tmp = 0;
  VBH_HIP(rocprim::radix_sort_keys(nullptr, tmp, in.p, out.p, m, 0u, static_cast<unsigned>(end_bit), s));
  DBuf<char> t;
  t.resize(std::max<size_t>(tmp, 1));
  VBH_HIP(rocprim::radix_sort_keys(t.p, tmp, in.p, out.p, m, 0u, static_cast<unsigned>(end_bit), s));
  VBH_HIP(hipStreamSynchronize(s));
}

// Stable radix sort of (key, position) pairs: equal keys keep their positions
// in increasing order.
template <typename V>
void sort_pairs(DBuf<unsigned long long>& kin, DBuf<unsigned long long>& kout, DBuf<V>& vin,
                DBuf<V>& vout, size_t m, int end_bit, hipStream_t s) {
  size_t tmp = 0;
  VBH_HIP(rocprim::radix_sort_pairs(nullptr, tmp, kin.p, kout.p, vin.p, vout.p, m, 0u,
                                    static_cast<unsigned>(end_bit), s));
  DBuf<char> t;
  t.resize(std::max<size_t>(tmp, 1));
  VBH_HIP(rocprim::radix_sort_pairs(t.p, tmp, kin.p, kout.p, vin.p, vout.p, m, 0u,
                                    static_cast<unsigned>(end_bit), s));
  VBH_HIP(hipStreamSynchronize(s));
}

void excl_scan(const int* in, int* out, size_t m, hipStream_t s) {
  size_t tmp = 0;
  VBH_HIP(rocprim::exclusive_scan(nullptr, tmp, in, out, 0, m, rocprim::plus<int>(), s));
  DBuf<char> t;
  t.resize(std::max<size_t>(tmp, 1));
  VBH_HIP(rocprim::exclusive_scan(t.p, tmp, in, out, 0, m, rocprim::plus<int>(), s));
  VBH_HIP(hipStreamSynchronize(s));
}

int read_last_plus(const int* scan, const int* flags, size_t m, hipStream_t s) {
  if (m == 0) return 0;
  int a = 0, b = 0;
  VBH_HIP(hipMemcpyAsync(&a, scan + m - 1, sizeof(int), hipMemcpyDeviceToHost, s));
  VBH_HIP(hipMemcpyAsync(&b, flags + m - 1, sizeof(int), hipMemcpyDeviceToHost, s));
  VBH_HIP(hipStreamSynchronize(s));
  return a + b;
}

}  // namespace

// Unique sorted (row,col) list from sorted keys: fills row_of/col, returns
// count; with seg, also each pair's first position in the sorted keys
// (seg[count] = position of the first sentinel).
static int unique_pairs(DBuf<unsigned long long>& keys, size_t m, int bits, DBuf<int>& row_of,
                        DBuf<int>& col, hipStream_t s, DBuf<long long>* seg = nullptr) {
  DBuf<int> head, pos;
  head.resize(std::max<size_t>(m, 1));
  pos.resize(std::max<size_t>(m, 1));
  const unsigned long long sentinel = (1ull << (2 * bits)) - 1ull;
  hipLaunchKernelGGL(k_mark_unique, dim3((m + 255) / 256), dim3(256), 0, s,
                     static_cast<long long>(m), keys.p, sentinel, head.p);
  excl_scan(head.p, pos.p, m, s);
  const int cnt = read_last_plus(pos.p, head.p, m, s);
  row_of.resize(std::max(cnt, 1));
  col.resize(std::max(cnt, 1));
  if (seg) seg->resize(static_cast<size_t>(cnt) + 1);
  hipLaunchKernelGGL(k_compact_keys, dim3((m + 255) / 256), dim3(256), 0, s,
                     static_cast<long long>(m), keys.p, head.p, pos.p, bits, row_of.p, col.p,
                     seg ? seg->p : nullptr);
  if (seg)
    hipLaunchKernelGGL(k_seg_end, dim3(1), dim3(1), 0, s, static_cast<long long>(m), keys.p,
                       sentinel, cnt, seg->p);
  VBH_HIP(hipStreamSynchronize(s));
  return cnt;
}

// Block pattern of the full symmetric operator of a lower CSR (entry rows
// row_of, columns ci): the sorted unique block pairs (brow, bcol).
struct BlockPattern {
  int B = 0, nbr = 0, nnzb = 0;
  DBuf<int> brow, bcol;
};

static void block_pattern(int n, int nnz, const int* row_of, const int* ci, int B, BlockPattern& p,
                          hipStream_t s) {
  p.B = B;
  p.nbr = (n + B - 1) / B;
  const int bits = bits_for(p.nbr);
  DBuf<unsigned long long> keys, sorted;
  const size_t m = 2 * static_cast<size_t>(nnz);
  keys.resize(m);
  sorted.resize(m);
  hipLaunchKernelGGL(k_block_keys, dim3((nnz + 255) / 256), dim3(256), 0, s, nnz, B, row_of, ci,
                     keys.p, bits);
  sort_keys(keys, sorted, m, 2 * bits, s);
  keys.release();
  p.nnzb = unique_pairs(sorted, m, bits, p.brow, p.bcol, s);
}

// Sorted distinct block-column offsets (bci - I) of the block rows each
// slice of kSliceRows rows overlaps: offsets of slice sl at
// offs[koff[sl] .. koff[sl + 1]).
void sell_slice_offsets(const BsrDev& A, int nslices, std::vector<int>& koff,
                        std::vector<int>& offs, hipStream_t s) {
  const int nbr = A.nbr;
  const int bits = bits_for(2ll * nbr);
  const unsigned long long sentinel = (1ull << (2 * bits)) - 1ull;
  const size_t m = 2 * static_cast<size_t>(A.nnzb);
  DBuf<unsigned long long> keys, sorted;
  keys.resize(std::max<size_t>(m, 1));
  sorted.resize(std::max<size_t>(m, 1));
  hipLaunchKernelGGL(k_rel_keys, dim3((nbr + 255) / 256), dim3(256), 0, s, nbr, A.B, A.brp, A.bci,
                     bits, sentinel, keys.p);
  sort_keys(keys, sorted, m, 2 * bits, s);
  keys.release();
  DBuf<int> slice_of, col, d_koff;
  const int cnt = unique_pairs(sorted, m, bits, slice_of, col, s);
  sorted.release();
  d_koff.resize(static_cast<size_t>(nslices) + 1);
  hipLaunchKernelGGL(k_row_ptr, dim3((cnt + 1 + 255) / 256), dim3(256), 0, s, cnt, nslices,
                     slice_of.p, d_koff.p);
  koff.resize(static_cast<size_t>(nslices) + 1);
  offs.resize(static_cast<size_t>(cnt));
  VBH_HIP(hipMemcpyAsync(koff.data(), d_koff.p, sizeof(int) * koff.size(), hipMemcpyDeviceToHost, s));
  if (cnt)
    VBH_HIP(hipMemcpyAsync(offs.data(), col.p, sizeof(int) * offs.size(), hipMemcpyDeviceToHost, s));
  VBH_HIP(hipStreamSynchronize(s));
  for (int& o : offs) o -= nbr;
}

// The 6x6 block pattern of the same matrix from its 3x3 one (p3.nnzb keys
// instead of two per lower entry).
static void coarsen_pattern(int n, const BlockPattern& p3, BlockPattern& p6, hipStream_t s) {
  p6.B = 6;
  p6.nbr = (n + 5) / 6;
  const int bits = bits_for(p6.nbr);
  const size_t m = static_cast<size_t>(p3.nnzb);
  DBuf<unsigned long long> keys, sorted;
  keys.resize(std::max<size_t>(m, 1));
  sorted.resize(std::max<size_t>(m, 1));
  hipLaunchKernelGGL(k_coarse_keys, dim3((p3.nnzb + 255) / 256), dim3(256), 0, s, p3.nnzb,
                     p3.brow.p, p3.bcol.p, bits, keys.p);
  sort_keys(keys, sorted, m, 2 * bits, s);
  keys.release();
  p6.nnzb = unique_pairs(sorted, m, bits, p6.brow, p6.bcol, s);
}

static double block_fill(const BlockPattern& p, long long full_nnz) {
  return static_cast<double>(p.nnzb) * p.B * p.B / static_cast<double>(full_nnz);
}

// BSR values of the device lower CSR held in ctx (l_row/l_ci/l_v) on pattern p.
static void fill_bsr(Ctx& c, BlockPattern& p) {
  hipStream_t s = c.stream;
  const int n = c.n, nnz = c.l_nnz, B = p.B, nbr = p.nbr, nnzb = p.nnzb;
  DBuf<int>& brow = p.brow;
  DBuf<int>& bcol = p.bcol;
  c.bsr_rp.resize(static_cast<size_t>(nbr) + 1);
  hipLaunchKernelGGL(k_row_ptr, dim3((nnzb + 1 + 255) / 256), dim3(256), 0, s, nnzb, nbr, brow.p,
                     c.bsr_rp.p);
  c.bsr_ci.resize(static_cast<size_t>(nnzb));
  VBH_HIP(hipMemcpyAsync(c.bsr_ci.p, bcol.p, sizeof(int) * nnzb, hipMemcpyDeviceToDevice, s));
  c.bsr_v64.resize(static_cast<size_t>(nnzb) * B * B);
  dzero(c.bsr_v64, s);
  if (B == 3) {
    hipLaunchKernelGGL(k_bsr_fill<3>, dim3((nnz + 255) / 256), dim3(256), 0, s, nnz, n, c.l_row.p,
                       c.l_ci.p, c.l_v.p, c.bsr_rp.p, c.bsr_ci.p, c.bsr_v64.p);
    if (nbr * 3 > n)
      hipLaunchKernelGGL(k_bsr_pad_diag<3>, dim3(1), dim3(64), 0, s, n, nbr, c.bsr_rp.p, c.bsr_ci.p,
                         c.bsr_v64.p);
  } else {
    hipLaunchKernelGGL(k_bsr_fill<6>, dim3((nnz + 255) / 256), dim3(256), 0, s, nnz, n, c.l_row.p,
                       c.l_ci.p, c.l_v.p, c.bsr_rp.p, c.bsr_ci.p, c.bsr_v64.p);
    if (nbr * 6 > n)
      hipLaunchKernelGGL(k_bsr_pad_diag<6>, dim3(1), dim3(64), 0, s, n, nbr, c.bsr_rp.p, c.bsr_ci.p,
                         c.bsr_v64.p);
  }
  c.bsr_v32.resize(c.bsr_v64.n);
  hipLaunchKernelGGL(k_cast, dim3(grid_for(c.bsr_v64.n, 1024)), dim3(256), 0, s,
                     static_cast<long long>(c.bsr_v64.n), c.bsr_v64.p, c.bsr_v32.p);
  VBH_HIP(hipStreamSynchronize(s));
  c.bsrd.B = B;
  c.bsrd.nbr = nbr;
  c.bsrd.nnzb = nnzb;
  c.bsrd.brp = c.bsr_rp.p;
  c.bsrd.bci = c.bsr_ci.p;
  c.bsrd.v64 = c.bsr_v64.p;
  c.bsrd.v32 = c.bsr_v32.p;
  build_sell(c);
}

// Build the BSR(B) operator on the device from the device lower CSR held in
// ctx (l_rp/l_ci/l_v/l_row). Returns false if the block fill is too high.
bool build_device_bsr(Ctx& c, int B) {
  BlockPattern p;
  block_pattern(c.n, c.l_nnz, c.l_row.p, c.l_ci.p, B, p, c.stream);
  // 6x6 blocks are requested for shell meshes, whose node blocks hold
  // structural zeros (fill about 1.2, the same as with 3x3 blocks)
  if (block_fill(p, 2ll * c.l_nnz - c.n) > (B == 6 ? 1.3 : 1.25)) return false;
  fill_bsr(c, p);
  return true;
}

void Ctx::assemble(int n_, int nnodes, const double* xyz_h, const int* eq_h, int nelem,
                   const int* etype_h, const int* eptr_h, const int* conn_h, const double* E_h,
                   const double* nu_h, const double* thick_h, int solid_form, int shell_form) {
  select_stream_device();
  hipStream_t s = stream;
  if (n_ <= 0) throw_error(VBH_ERR_SOLVER, "HIP PCG: stiffness matrix is empty -- no free DOFs");
  if (nelem <= 0 || nnodes <= 0) throw_error(VBH_ERR_ARG, "vbh_assemble: empty mesh");
  if (mpc.active && mpc.n_full != n_)
    throw_error(VBH_ERR_ARG, "vbh_assemble: n differs from the n_full of vbh_set_mpc");
  // host-side grouping by element kind
  constexpr int kQuad4 = 4;  // group index of CQUAD4 shells
  std::vector<int> groups[5];
  for (int e = 0; e < nelem; ++e) {
    const int t = etype_h[e], nn = eptr_h[e + 1] - eptr_h[e];
    if (t == VBH_CHEXA8 && nn == 8) groups[solid_form == VBH_SOLID_SRI ? kHexSri : kHexEas].push_back(e);
    else if (t == VBH_CPENTA6 && nn == 6) groups[kPentaEas].push_back(e);
    else if (t == VBH_CTETRA4 && nn == 4) groups[kTet4].push_back(e);
    else if (t == VBH_CQUAD4 && nn == 4) groups[kQuad4].push_back(e);
    else throw_error(VBH_ERR_ARG, "vbh_assemble: unsupported element type or node count at element " +
                                      std::to_string(e));
  }
  if (!groups[kQuad4].empty()) {
    if (!thick_h) throw_error(VBH_ERR_ARG, "vbh_assemble: CQUAD4 elements need a thickness array");
    if (shell_form != VBH_SHELL_MITC4 && shell_form != VBH_SHELL_MINDLIN)
      throw_error(VBH_ERR_ARG, "vbh_assemble: unknown shell formulation");
    for (int e : groups[kQuad4])
      if (!(thick_h[e] > 0.0))
        throw_error(VBH_ERR_ARG, "vbh_assemble: CQUAD4 element " + std::to_string(e) +
                                     " has non-positive thickness");
  }
  const int nconn = eptr_h[nelem];
  for (int k = 0; k < nconn; ++k)
    if (conn_h[k] < 0 || conn_h[k] >= nnodes)
      throw_error(VBH_ERR_ARG, "vbh_assemble: connectivity index out of range");
  for (int k = 0; k < 6 * nnodes; ++k)
    if (eq_h[k] < -1 || eq_h[k] >= n_) throw_error(VBH_ERR_ARG, "vbh_assemble: eq index out of range");

  DBuf<double> d_xyz, d_E, d_nu, d_t;
  DBuf<int> d_eq, d_et, d_ep, d_conn, d_gd, d_err;
  d_xyz.upload(xyz_h, static_cast<size_t>(nnodes) * 3, s);
  d_eq.upload(eq_h, static_cast<size_t>(nnodes) * 6, s);
  d_et.upload(etype_h, static_cast<size_t>(nelem), s);
  d_ep.upload(eptr_h, static_cast<size_t>(nelem) + 1, s);
  d_conn.upload(conn_h, static_cast<size_t>(nconn), s);
  d_E.upload(E_h, static_cast<size_t>(nelem), s);
  d_nu.upload(nu_h, static_cast<size_t>(nelem), s);
  if (!groups[kQuad4].empty()) d_t.upload(thick_h, static_cast<size_t>(nelem), s);
  d_gd.resize(static_cast<size_t>(nelem) * kGdStride);
  d_err.resize(4);
  hipLaunchKernelGGL(k_init_err, dim3(1), dim3(1), 0, s, d_err.p);
  hipLaunchKernelGGL(k_gdofs, dim3((nelem + 255) / 256), dim3(256), 0, s, nelem, d_et.p, d_ep.p,
                     d_conn.p, d_eq.p, d_gd.p);
  DBuf<int> d_sel[5];
  for (int k = 0; k < 5; ++k)
    if (!groups[k].empty()) d_sel[k].upload(groups[k].data(), groups[k].size(), s);

  hipEvent_t e0, e1, e2;
  VBH_HIP(hipEventCreate(&e0));
  VBH_HIP(hipEventCreate(&e1));
  VBH_HIP(hipEventCreate(&e2));
  VBH_HIP(hipEventRecord(e0, s));

  // 2. topological pattern
  const int bits = bits_for(n_);
  size_t m = 0;
  const int ndk[5] = {24, 24, 18, 12, 24};
  size_t off_k[5] = {0, 0, 0, 0, 0};
  for (int k = 0; k < 5; ++k) {
    off_k[k] = m;
    m += groups[k].size() * static_cast<size_t>(ndk[k] * (ndk[k] + 1) / 2);
  }
  if (m >= (1ull << 32))
    throw_error(VBH_ERR_ARG, "vbh_assemble: more than 2^32 element matrix entries");
  int ngroups = 0;
  for (int k = 0; k < 5; ++k) ngroups += !groups[k].empty();
  int nnz_topo = 0;
  DBuf<int> t_row, t_col;
  // node-level pattern (pass 2') when every element carries the same DOFs per
  // node: D = 3 (solids only) or 6 (CQUAD4 only); VBH_PATTERN=dof takes the
  // pair-level pass instead
  int D = 0;
  {
    const bool shells = !groups[kQuad4].empty();
    const bool solids = ngroups > (shells ? 1 : 0);
    const char* pe = std::getenv("VBH_PATTERN");
    const bool force_pairs = pe && std::strcmp(pe, "dof") == 0;
    if (!force_pairs && shells != solids && nelem < (1 << 26)) D = shells ? 6 : 3;
  }
  DBuf<int> np_rb, np_cb, np_pre, np_first, np_rowptr, np_node_of_eq, np_info, np_off;
  DBuf<long long> np_seg;
  DBuf<unsigned long long> np_vals;
  DBuf<unsigned> np_ebase;
  int np_u = 0;
  bool node_level = false;
  if (D) {
    std::vector<int> npo(static_cast<size_t>(nelem) + 1, 0);
    long long tot = 0;
    for (int e = 0; e < nelem; ++e) {
      const int nn = eptr_h[e + 1] - eptr_h[e];
      npo[e] = static_cast<int>(tot);
      tot += nn * (nn + 1) / 2;
    }
    if (tot < (1ll << 31)) {
      npo[nelem] = static_cast<int>(tot);
      std::vector<unsigned> eb(static_cast<size_t>(nelem), 0u);
      for (int k = 0; k < 5; ++k) {
        const unsigned nl = static_cast<unsigned>(ndk[k] * (ndk[k] + 1) / 2);
        for (size_t q = 0; q < groups[k].size(); ++q)
          eb[groups[k][q]] = static_cast<unsigned>(off_k[k]) + static_cast<unsigned>(q) * nl;
      }
      DBuf<int> d_npo, nbase, nact, dmask, bad;
      d_npo.upload(npo.data(), npo.size(), s);
      np_ebase.upload(eb.data(), eb.size(), s);
      nbase.resize(static_cast<size_t>(nnodes));
      nact.resize(static_cast<size_t>(nnodes));
      dmask.resize(static_cast<size_t>(nnodes));
      np_node_of_eq.resize(static_cast<size_t>(n_));
      hipLaunchKernelGGL(k_set_i32, dim3(grid_for(n_, 1024)), dim3(256), 0, s,
                         static_cast<long long>(n_), np_node_of_eq.p, -1);
      bad.resize(1);
      dzero(bad, s);
      hipLaunchKernelGGL(k_node_dofs, dim3((nnodes + 255) / 256), dim3(256), 0, s, nnodes, D,
                         d_eq.p, nbase.p, nact.p, dmask.p, np_node_of_eq.p, bad.p);
      const size_t mn = static_cast<size_t>(tot);
      DBuf<unsigned long long> keys, sorted, vals;
      keys.resize(std::max<size_t>(mn, 1));
      sorted.resize(std::max<size_t>(mn, 1));
      vals.resize(std::max<size_t>(mn, 1));
      np_vals.resize(std::max<size_t>(mn, 1));
      hipLaunchKernelGGL(k_node_pair_keys, dim3((nelem + 255) / 256), dim3(256), 0, s, nelem,
                         d_ep.p, d_conn.p, d_npo.p, nbase.p, nact.p, np_ebase.p, bits, keys.p,
                         vals.p, bad.p);
      int hbad = 0;
      VBH_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
      VBH_HIP(hipStreamSynchronize(s));
      if (!hbad && mn > 0) {
        sort_pairs(keys, sorted, vals, np_vals, mn, 2 * bits, s);
        keys.release();
        vals.release();
        np_u = unique_pairs(sorted, mn, bits, np_rb, np_cb, s, &np_seg);
        sorted.release();
        DBuf<int> cnt, diag, len;
        cnt.resize(static_cast<size_t>(std::max(np_u, 1)));
        np_info.resize(static_cast<size_t>(std::max(np_u, 1)));
        np_off.resize(static_cast<size_t>(std::max(np_u, 1)));
        np_first.resize(static_cast<size_t>(nnodes));
        diag.resize(static_cast<size_t>(nnodes));
        hipLaunchKernelGGL(k_set_i32, dim3(grid_for(nnodes, 1024)), dim3(256), 0, s,
                           static_cast<long long>(nnodes), diag.p, -1);
        if (np_u > 0)
          hipLaunchKernelGGL(k_np_info, dim3((np_u + 255) / 256), dim3(256), 0, s, np_u, np_rb.p,
                             np_cb.p, np_node_of_eq.p, nact.p, dmask.p, cnt.p, np_first.p, diag.p,
                             np_info.p);
        np_pre.resize(static_cast<size_t>(std::max(np_u, 1)));
        excl_scan(cnt.p, np_pre.p, static_cast<size_t>(np_u), s);
        if (np_u > 0)
          hipLaunchKernelGGL(k_np_off, dim3((np_u + 255) / 256), dim3(256), 0, s, np_u, np_rb.p,
                             np_node_of_eq.p, np_pre.p, np_first.p, np_off.p);
        len.resize(static_cast<size_t>(n_));
        hipLaunchKernelGGL(k_np_row_len, dim3((n_ + 255) / 256), dim3(256), 0, s, n_,
                           np_node_of_eq.p, nbase.p, np_pre.p, np_first.p, diag.p, len.p);
        np_rowptr.resize(static_cast<size_t>(n_));
        excl_scan(len.p, np_rowptr.p, static_cast<size_t>(n_), s);
        nnz_topo = read_last_plus(np_rowptr.p, len.p, static_cast<size_t>(n_), s);
        node_level = true;
      }
    }
  }
  DBuf<long long> seg;       // slot p's run in the sorted pair list
  DBuf<unsigned> src_sorted;  // pair positions in sorted key order
  DBuf<int> eid;             // element of each pair position (several kinds only)
  if (!node_level) {
    DBuf<unsigned long long> keys, sorted;
    DBuf<unsigned> src;
    keys.resize(m);
    sorted.resize(m);
    src.resize(m);
    src_sorted.resize(m);
    if (ngroups > 1) eid.resize(m);
    for (int k = 0; k < 5; ++k) {
      const int ns = static_cast<int>(groups[k].size());
      if (!ns) continue;
      const int NL = ndk[k] * (ndk[k] + 1) / 2;
      const long long cnt = static_cast<long long>(ns) * NL;
      const dim3 g(static_cast<unsigned>((cnt + 255) / 256));
      const size_t off = off_k[k];
      const unsigned o32 = static_cast<unsigned>(off);
      if (ndk[k] == 24)
        hipLaunchKernelGGL(k_pair_keys<24>, g, dim3(256), 0, s, ns, d_sel[k].p, d_gd.p,
                           keys.p + off, bits, src.p + off, o32, eid.p ? eid.p + off : nullptr);
      else if (ndk[k] == 18)
        hipLaunchKernelGGL(k_pair_keys<18>, g, dim3(256), 0, s, ns, d_sel[k].p, d_gd.p,
                           keys.p + off, bits, src.p + off, o32, eid.p ? eid.p + off : nullptr);
      else
        hipLaunchKernelGGL(k_pair_keys<12>, g, dim3(256), 0, s, ns, d_sel[k].p, d_gd.p,
                           keys.p + off, bits, src.p + off, o32, eid.p ? eid.p + off : nullptr);
    }
    sort_pairs(keys, sorted, src, src_sorted, m, 2 * bits, s);
    keys.release();
    src.release();
    nnz_topo = unique_pairs(sorted, m, bits, t_row, t_col, s, &seg);
  }
  VBH_HIP(hipEventRecord(e1, s));

  // 3. element matrices (entries at their pair positions), then slot sums
  DBuf<double> ke;
  ke.resize(m);
  for (int k = 0; k < 5; ++k) {
    const int ns = static_cast<int>(groups[k].size());
    if (!ns) continue;
    double* ke_k = ke.p + off_k[k];
    if (k == kQuad4) {
      launch_quad4_stiffness(ns, d_sel[k].p, d_ep.p, d_conn.p, d_xyz.p, d_E.p, d_nu.p, d_t.p,
                             shell_form, d_gd.p, ke_k, d_err.p, s);
      continue;
    }
    const dim3 g(static_cast<unsigned>(std::min((ns + kElemWaves - 1) / kElemWaves, 65535)));
    const dim3 b(64 * kElemWaves);
    if (k == kHexEas)
      hipLaunchKernelGGL(k_element_stiffness<kHexEas>, g, b, 0, s, ns, d_sel[k].p, d_ep.p, d_conn.p,
                         d_xyz.p, d_E.p, d_nu.p, d_gd.p, nullptr, nullptr, ke_k, nullptr, d_err.p);
    else if (k == kHexSri)
      hipLaunchKernelGGL(k_element_stiffness<kHexSri>, g, b, 0, s, ns, d_sel[k].p, d_ep.p, d_conn.p,
                         d_xyz.p, d_E.p, d_nu.p, d_gd.p, nullptr, nullptr, ke_k, nullptr, d_err.p);
    else if (k == kPentaEas)
      hipLaunchKernelGGL(k_element_stiffness<kPentaEas>, g, b, 0, s, ns, d_sel[k].p, d_ep.p,
                         d_conn.p, d_xyz.p, d_E.p, d_nu.p, d_gd.p, nullptr, nullptr, ke_k, nullptr,
                         d_err.p);
    else
      hipLaunchKernelGGL(k_element_stiffness<kTet4>, g, b, 0, s, ns, d_sel[k].p, d_ep.p, d_conn.p,
                         d_xyz.p, d_E.p, d_nu.p, d_gd.p, nullptr, nullptr, ke_k, nullptr, d_err.p);
  }
  DBuf<double> t_val;
  DBuf<int> touched;
  t_val.resize(static_cast<size_t>(std::max(nnz_topo, 1)));
  touched.resize(static_cast<size_t>(std::max(nnz_topo, 1)));
  if (node_level) {
    t_row.resize(static_cast<size_t>(std::max(nnz_topo, 1)));
    t_col.resize(static_cast<size_t>(std::max(nnz_topo, 1)));
    if (np_u > 0) {
      const dim3 g(static_cast<unsigned>((static_cast<long long>(np_u) * D * D + 255) / 256));
      if (D == 3)
        hipLaunchKernelGGL(k_np_slot_sums<3>, g, dim3(256), 0, s, np_u, np_rb.p, np_cb.p, np_seg.p,
                           np_vals.p, np_info.p, np_off.p, np_rowptr.p, ke.p, t_row.p, t_col.p,
                           t_val.p, touched.p);
      else
        hipLaunchKernelGGL(k_np_slot_sums<6>, g, dim3(256), 0, s, np_u, np_rb.p, np_cb.p, np_seg.p,
                           np_vals.p, np_info.p, np_off.p, np_rowptr.p, ke.p, t_row.p, t_col.p,
                           t_val.p, touched.p);
    }
  } else if (nnz_topo > 0 && ngroups > 1)
    hipLaunchKernelGGL(k_slot_sums_by_element, dim3((nnz_topo + 255) / 256), dim3(256), 0, s,
                       nnz_topo, seg.p, src_sorted.p, eid.p, ke.p, t_val.p, touched.p);
  else if (nnz_topo > 0)
    hipLaunchKernelGGL(k_slot_sums, dim3((nnz_topo + 255) / 256), dim3(256), 0, s, nnz_topo, seg.p,
                       src_sorted.p, ke.p, t_val.p, touched.p);
  VBH_HIP(hipEventRecord(e2, s));
  int herr[4];
  VBH_HIP(hipMemcpyAsync(herr, d_err.p, sizeof herr, hipMemcpyDeviceToHost, s));
  VBH_HIP(hipStreamSynchronize(s));
  float ms_pat = 0, ms_elem = 0;
  VBH_HIP(hipEventElapsedTime(&ms_pat, e0, e1));
  VBH_HIP(hipEventElapsedTime(&ms_elem, e1, e2));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  const std::string bad_e = "element " + std::to_string(herr[1]);
  if (herr[0] & kErrJacobian) throw_error(VBH_ERR_SOLVER, bad_e + ": non-positive Jacobian");
  if (herr[0] & kErrKaa) throw_error(VBH_ERR_SOLVER, "EAS element: Kaa factorization failed");
  if (herr[0] & kErrMitcGeometry)
    throw_error(VBH_ERR_SOLVER, "CQUAD4 " + bad_e + ": MITC4+ characteristic geometry is singular");
  if (herr[0] & kErrMitcDistort)
    throw_error(VBH_ERR_SOLVER, "CQUAD4 " + bad_e + ": MITC4+ distortion denominator is near zero");
  if (herr[0] & kErrShellDiag)
    throw_error(VBH_ERR_SOLVER, "CQUAD4 " + bad_e +
                                    ": stiffness matrix diagonal is near-zero; check material, "
                                    "thickness, and geometry");
  if (herr[0] & kErrSlot) throw_error(VBH_ERR_HIP, "assembly: CSR slot not found (internal error)");

  // 4. compaction of slots without a nonzero contribution
  DBuf<int> pos;
  pos.resize(std::max(nnz_topo, 1));
  excl_scan(touched.p, pos.p, static_cast<size_t>(nnz_topo), s);
  const int nnz = read_last_plus(pos.p, touched.p, static_cast<size_t>(nnz_topo), s);
  l_ci.resize(static_cast<size_t>(std::max(nnz, 1)));
  l_v.resize(static_cast<size_t>(std::max(nnz, 1)));
  l_row.resize(static_cast<size_t>(std::max(nnz, 1)));
  hipLaunchKernelGGL(k_compact_csr, dim3((nnz_topo + 255) / 256), dim3(256), 0, s, nnz_topo,
                     touched.p, pos.p, t_col.p, t_val.p, t_row.p, l_ci.p, l_v.p, l_row.p);
  l_rp.resize(static_cast<size_t>(n_) + 1);
  hipLaunchKernelGGL(k_row_ptr, dim3((nnz + 1 + 255) / 256), dim3(256), 0, s, nnz, n_, l_row.p,
                     l_rp.p);
  VBH_HIP(hipStreamSynchronize(s));
  n = n_;
  l_nnz = nnz;
  asm_stats.nnz_topological = nnz_topo;
  asm_stats.nnz = nnz;
  asm_stats.pattern_ms = ms_pat;
  asm_stats.element_ms = ms_elem;
  lower_on_host = false;
  {
    bool shells = false, solids = false;
    for (int e = 0; e < nelem; ++e) (etype_h[e] == VBH_CQUAD4 ? shells : solids) = true;
    block_hint = (shells && !solids) ? 6 : (solids && !shells ? 3 : 0);
  }
  // keep the mesh on the device for the sensitivity kernels
  mesh.valid = false;
  mesh.xyz.swap(d_xyz);
  mesh.E.swap(d_E);
  mesh.nu.swap(d_nu);
  mesh.t.swap(d_t);
  mesh.eq.swap(d_eq);
  mesh.et.swap(d_et);
  mesh.ep.swap(d_ep);
  mesh.conn.swap(d_conn);
  mesh.gd.swap(d_gd);
  for (int k = 0; k < 5; ++k) {
    mesh.sel[k].swap(d_sel[k]);
    mesh.nsel[k] = static_cast<int>(groups[k].size());
  }
  mesh.n = n_;
  mesh.nnodes = nnodes;
  mesh.nelem = nelem;
  mesh.solid_form = solid_form;
  mesh.shell_form = shell_form;
  asm_stats.mpc_ms = 0.0;
  op_padded = false;
  n_api = n_;
  if (mpc.active) {
    // K_red = T^T K_full T in the full numbering (unit diagonal on dependents)
    const auto t0 = std::chrono::steady_clock::now();
    mesh.mpc.upload(mpc, n_, s);
    apply_mpc_to_lower();
    asm_stats.mpc_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    asm_stats.nnz = l_nnz - mesh.mpc.n_dep;
  } else {
    mesh.mpc.upload(mpc, n_, s);
  }
  mesh.valid = true;
  finish_device_operator();
}

// out[e] = -u_e^T (dK_e/dp_e) v_e over the retained mesh (param 0: E, 1: CQUAD4
// thickness). Host or device vectors.
void Ctx::element_sensitivity(int param, const double* u, const double* v, double* out, bool dev) {
  if (!mesh.valid)
    throw_error(VBH_ERR_STATE, "vbh_element_sensitivity: no assembled mesh in this context");
  if (param != 0 && param != 1)
    throw_error(VBH_ERR_ARG, "vbh_element_sensitivity: param must be 0 (E) or 1 (thickness)");
  if (!u || !v || !out) throw_error(VBH_ERR_ARG, "vbh_element_sensitivity: null vector");
  select_stream_device();
  hipStream_t s = stream;
  DBuf<double> du, dv, dout, fu, fv;
  const double* pu = mesh_solution(u, dev, du, fu);
  const double* pv = mesh_solution(v, dev, dv, fv);
  double* po = out;
  if (!dev) {
    dout.resize(static_cast<size_t>(mesh.nelem));
    po = dout.p;
  }
  launch_fill_f64(mesh.nelem, po, 0.0, s);  // solids stay 0 for thickness
  DBuf<int> d_err;
  d_err.resize(4);
  hipLaunchKernelGGL(k_init_err, dim3(1), dim3(1), 0, s, d_err.p);
  for (int k = 0; k < 5; ++k) {
    const int ns = mesh.nsel[k];
    if (!ns) continue;
    if (k == 4) {
      launch_quad4_sensitivity(ns, mesh.sel[k].p, mesh.ep.p, mesh.conn.p, mesh.xyz.p, mesh.E.p,
                               mesh.nu.p, mesh.t.p, mesh.shell_form, mesh.gd.p, param, pu, pv, po,
                               d_err.p, s);
      continue;
    }
    if (param != 0) continue;
    const dim3 g(static_cast<unsigned>(std::min((ns + kElemWaves - 1) / kElemWaves, 65535)));
    const dim3 b(64 * kElemWaves);
#define VBH_SENS(KIND)                                                                          \
  hipLaunchKernelGGL((k_element_stiffness<KIND, 1>), g, b, 0, s, ns, mesh.sel[k].p, mesh.ep.p,  \
                     mesh.conn.p, mesh.xyz.p, mesh.E.p, mesh.nu.p, mesh.gd.p, nullptr, nullptr, \
                     nullptr, nullptr, d_err.p, pu, pv, po)
    if (k == kHexEas) VBH_SENS(kHexEas);
    else if (k == kHexSri) VBH_SENS(kHexSri);
    else if (k == kPentaEas) VBH_SENS(kPentaEas);
    else VBH_SENS(kTet4);
#undef VBH_SENS
  }
  int herr[4];
  VBH_HIP(hipMemcpyAsync(herr, d_err.p, sizeof herr, hipMemcpyDeviceToHost, s));
  if (!dev)
    VBH_HIP(hipMemcpyAsync(out, dout.p, sizeof(double) * mesh.nelem, hipMemcpyDeviceToHost, s));
  VBH_HIP(hipStreamSynchronize(s));
  if (herr[0]) throw_error(VBH_ERR_SOLVER, "vbh_element_sensitivity: element " +
                                               std::to_string(herr[1]) + " failed (code " +
                                               std::to_string(herr[0]) + ")");
}

// f[6 node + dof] = sum_e K_e u_e over the retained mesh (internal nodal
// forces at every DOF, constrained ones included). Host or device vectors.
void Ctx::element_forces(const double* u, double* fout, bool dev) {
  if (!mesh.valid) throw_error(VBH_ERR_STATE, "vbh_element_forces: no assembled mesh in this context");
  if (!u || !fout) throw_error(VBH_ERR_ARG, "vbh_element_forces: null vector");
  select_stream_device();
  hipStream_t s = stream;
  const size_t m = 6 * static_cast<size_t>(mesh.nnodes);
  DBuf<double> du, dout, fu;
  const double* pu = mesh_solution(u, dev, du, fu);
  double* po = fout;
  if (!dev) {
    dout.resize(m);
    po = dout.p;
  }
  launch_fill_f64(static_cast<long long>(m), po, 0.0, s);
  DBuf<int> d_err;
  d_err.resize(4);
  hipLaunchKernelGGL(k_init_err, dim3(1), dim3(1), 0, s, d_err.p);
  for (int k = 0; k < 5; ++k) {
    const int ns = mesh.nsel[k];
    if (!ns) continue;
    if (k == 4) {
      launch_quad4_forces(ns, mesh.sel[k].p, mesh.ep.p, mesh.conn.p, mesh.xyz.p, mesh.E.p,
                          mesh.nu.p, mesh.t.p, mesh.shell_form, mesh.gd.p, pu, po, d_err.p, s);
      continue;
    }
    const dim3 g(static_cast<unsigned>(std::min((ns + kElemWaves - 1) / kElemWaves, 65535)));
    const dim3 b(64 * kElemWaves);
#define VBH_FORCES(KIND)                                                                        \
  hipLaunchKernelGGL((k_element_stiffness<KIND, 2>), g, b, 0, s, ns, mesh.sel[k].p, mesh.ep.p,  \
                     mesh.conn.p, mesh.xyz.p, mesh.E.p, mesh.nu.p, mesh.gd.p, nullptr, nullptr, \
                     nullptr, nullptr, d_err.p, pu, nullptr, po)
    if (k == kHexEas) VBH_FORCES(kHexEas);
    else if (k == kHexSri) VBH_FORCES(kHexSri);
    else if (k == kPentaEas) VBH_FORCES(kPentaEas);
    else VBH_FORCES(kTet4);
#undef VBH_FORCES
  }
  VBH_HIP(hipGetLastError());
  int herr[4];
  VBH_HIP(hipMemcpyAsync(herr, d_err.p, sizeof herr, hipMemcpyDeviceToHost, s));
  if (!dev) VBH_HIP(hipMemcpyAsync(fout, dout.p, sizeof(double) * m, hipMemcpyDeviceToHost, s));
  VBH_HIP(hipStreamSynchronize(s));
  if (herr[0]) throw_error(VBH_ERR_SOLVER, "vbh_element_forces: element " +
                                               std::to_string(herr[1]) + " failed (code " +
                                               std::to_string(herr[0]) + ")");
}

// Device operator from the device lower CSR: diagonal checks, BSR build (or a
// host round trip for matrices without block structure).
void Ctx::finish_device_operator() {
  if (2ll * l_nnz >= 0x7fffffffll)
    throw_error(VBH_ERR_ARG, "assembly: 2^30 or more lower entries (the block build counts its "
                             "keys with 32-bit indices)");
  // Block size: 6 when all rows come in node groups of six, else 3.
  int B = format_req;
  if (B < 0) B = block_hint;
  bool ok = false;
  if (B == 3 || B == 6) ok = build_device_bsr(*this, B);
  if (!ok) {
    // no usable block structure: go through the host conversion path
    fetch_lower_to_host();
    const std::vector<int> rp = lower_rp, ci = lower_ci;
    const std::vector<double> v = lower_v;
    upload_operator(n, static_cast<int>(ci.size()), rp.data(), ci.data(), v.data(), true);
    return;
  }
  finish_block_operator(B);
}

// Jacobi diagonal, checks and bookkeeping of a BSR(B) operator built on the
// device from the device lower CSR.
void Ctx::finish_block_operator(int B) {
  hipStream_t s = stream;
  format = B;
  n_pad = bsrd.nbr * B;
  DBuf<int> d_err;
  d_err.resize(4);
  hipLaunchKernelGGL(k_init_err, dim3(1), dim3(1), 0, s, d_err.p);
  diag64.resize(static_cast<size_t>(n_pad));
  diag32.resize(static_cast<size_t>(n_pad));
  hipLaunchKernelGGL(k_diag_from_lower, dim3((n_pad + 255) / 256), dim3(256), 0, s, n, n_pad,
                     l_rp.p, l_ci.p, l_v.p, diag64.p, diag32.p, d_err.p);
  int herr[4];
  VBH_HIP(hipMemcpyAsync(herr, d_err.p, sizeof herr, hipMemcpyDeviceToHost, s));
  VBH_HIP(hipStreamSynchronize(s));
  if (herr[0] & 16)
    throw_error(VBH_ERR_SOLVER, "HIP PCG: missing diagonal at row " + std::to_string(herr[2]) +
                                    " -- matrix is singular. Check boundary conditions.");
  if (herr[0] & 8)
    throw_error(VBH_ERR_SOLVER, "HIP PCG: zero diagonal at row " + std::to_string(herr[2]) +
                                    " -- matrix is singular. Check boundary conditions.");
  // scalar CSR is not built for device-assembled operators
  csr = CsrDev{};
  csr.n = n;
  csr_rp.release();
  csr_ci.release();
  csr_v64.release();
  csr_v32.release();
  operator_version += 1;
  have_matrix = true;
}

void Ctx::fetch_lower_to_host() {
  if (lower_on_host) return;
  lower_rp.resize(static_cast<size_t>(n) + 1);
  lower_ci.resize(static_cast<size_t>(l_nnz));
  lower_v.resize(static_cast<size_t>(l_nnz));
  VBH_HIP(hipMemcpyAsync(lower_rp.data(), l_rp.p, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, stream));
  if (l_nnz) {
    VBH_HIP(hipMemcpyAsync(lower_ci.data(), l_ci.p, sizeof(int) * l_nnz, hipMemcpyDeviceToHost, stream));
    VBH_HIP(hipMemcpyAsync(lower_v.data(), l_v.p, sizeof(double) * l_nnz, hipMemcpyDeviceToHost, stream));
  }
  VBH_HIP(hipStreamSynchronize(stream));
  lower_on_host = true;
}

// Operator from a caller's lower CSR (vbh_set_csr) built on the device: the
// arrays are uploaded as they are, checked row by row, and blocked by the
// same rule as the host conversion. Returns false, with the context
// untouched, when the rows are not strictly increasing or no block format
// fits, or when the matrix has 2^30 entries or more (the host conversion then
// handles the matrix); throws the host conversion's errors.
bool Ctx::set_lower_device(int n_, int nnz, const int* rp, const int* ci, const double* v) {
  hipStream_t s = stream;
  // the block keys (two per entry) are counted with 32-bit scans
  if (2ll * nnz >= 0x7fffffffll) return false;
  DBuf<int> t_rp, t_ci, t_row;
  DBuf<double> t_v;
  t_rp.upload(rp, static_cast<size_t>(n_) + 1, s);
  t_ci.upload(ci, static_cast<size_t>(nnz), s);
  t_v.upload(v, static_cast<size_t>(nnz), s);
  t_row.resize(static_cast<size_t>(std::max(nnz, 1)));
  struct Check {
    unsigned long long first_err;
    int unsorted;
  };
  DBuf<Check> d_chk;
  d_chk.resize(1);
  hipLaunchKernelGGL(k_init_check, dim3(1), dim3(1), 0, s, &d_chk.p->first_err, &d_chk.p->unsorted);
  hipLaunchKernelGGL(k_check_lower_rows, dim3((n_ + 255) / 256), dim3(256), 0, s, n_, t_rp.p,
                     t_ci.p, t_v.p, t_row.p, &d_chk.p->first_err, &d_chk.p->unsorted);
  Check h{};
  VBH_HIP(hipMemcpyAsync(&h, d_chk.p, sizeof h, hipMemcpyDeviceToHost, s));
  VBH_HIP(hipStreamSynchronize(s));
  if (h.first_err != ~0ull) {
    const std::string row = std::to_string(h.first_err >> 3);
    switch (static_cast<int>(h.first_err & 7)) {
      case 1: throw_error(VBH_ERR_ARG, "vbh_set_csr: column index out of range");
      case 2: throw_error(VBH_ERR_ARG, "vbh_set_csr: upper entry in lower storage");
      case 3:
        throw_error(VBH_ERR_SOLVER, "HIP PCG: zero diagonal at row " + row +
                                        " -- matrix is singular. Check boundary conditions.");
      default:
        throw_error(VBH_ERR_SOLVER, "HIP PCG: missing diagonal at row " + row +
                                        " -- matrix is singular. Check boundary conditions.");
    }
  }
  if (h.unsorted) return false;
  // Block size: the host conversion's rule (6x6 node blocks when they carry
  // no more padding than 3x3 blocks would, else 3x3 when the fill is low).
  const long long full_nnz = 2ll * nnz - n_;
  BlockPattern p6, p3;
  BlockPattern* pick = nullptr;
  if (format_req == 3 || format_req == 6) {
    BlockPattern& p = format_req == 6 ? p6 : p3;
    block_pattern(n_, nnz, t_row.p, t_ci.p, format_req, p, s);
    pick = &p;
  } else {
    // (the 6x6 pattern is derived from the 3x3 one, which is built first)
    double f6 = 1e300, f3 = 1e300;
    if (n_ >= 12) {
      block_pattern(n_, nnz, t_row.p, t_ci.p, 3, p3, s);
      if (n_ >= 24) {
        coarsen_pattern(n_, p3, p6, s);
        f6 = block_fill(p6, full_nnz);
      }
      if (f6 > 1.12) f3 = block_fill(p3, full_nnz);
    }
    if (f6 <= 1.12 || (f6 <= 1.3 && f6 <= 1.05 * f3)) pick = &p6;
    else if (f3 <= 1.25) pick = &p3;
    else return false;
  }
  n = n_;
  l_nnz = nnz;
  l_rp.swap(t_rp);
  l_ci.swap(t_ci);
  l_v.swap(t_v);
  l_row.swap(t_row);
  lower_on_host = false;
  lower_rp.clear();
  lower_ci.clear();
  lower_v.clear();
  fill_bsr(*this, *pick);
  finish_block_operator(pick->B);
  return true;
}

// GPU DOF numbering (closed form of build_analysis_dof_map).
int Ctx::dof_map(int nnodes, const int* ids_h, int nelem, const int* etype_h, const int* eptr_h,
                 const int* conn_h, int nspc, const int* spc_nodes_h, const int* spc_masks_h,
                 int* eq_out) {
  select_stream_device();
  hipStream_t s = stream;
  if (nnodes <= 0) throw_error(VBH_ERR_ARG, "vbh_dof_map: no nodes");
  const int nconn = nelem > 0 ? eptr_h[nelem] : 0;
  for (int k = 0; k < nconn; ++k)
    if (conn_h[k] < 0 || conn_h[k] >= nnodes) throw_error(VBH_ERR_ARG, "vbh_dof_map: bad connectivity");
  for (int k = 0; k < nspc; ++k)
    if (spc_nodes_h[k] < 0 || spc_nodes_h[k] >= nnodes) throw_error(VBH_ERR_ARG, "vbh_dof_map: bad SPC node");
  DBuf<int> ids, idx, ids_s, order, et, ep, cn, mask, spc, sn, sm, cnt, base, eq;
  ids.upload(ids_h, static_cast<size_t>(nnodes), s);
  std::vector<int> iota(static_cast<size_t>(nnodes));
  std::iota(iota.begin(), iota.end(), 0);
  idx.upload(iota.data(), iota.size(), s);
  ids_s.resize(static_cast<size_t>(nnodes));
  order.resize(static_cast<size_t>(nnodes));
  {
    size_t tmp = 0;
    VBH_HIP(rocprim::radix_sort_pairs(nullptr, tmp, ids.p, ids_s.p, idx.p, order.p,
                                      static_cast<size_t>(nnodes), 0u, 32u, s));
    DBuf<char> t;
    t.resize(std::max<size_t>(tmp, 1));
    VBH_HIP(rocprim::radix_sort_pairs(t.p, tmp, ids.p, ids_s.p, idx.p, order.p,
                                      static_cast<size_t>(nnodes), 0u, 32u, s));
    VBH_HIP(hipStreamSynchronize(s));
  }
  mask.resize(static_cast<size_t>(nnodes));
  spc.resize(static_cast<size_t>(nnodes));
  dzero(mask, s);
  dzero(spc, s);
  if (nelem > 0) {
    et.upload(etype_h, static_cast<size_t>(nelem), s);
    ep.upload(eptr_h, static_cast<size_t>(nelem) + 1, s);
    cn.upload(conn_h, static_cast<size_t>(std::max(nconn, 1)), s);
    hipLaunchKernelGGL(k_dof_active, dim3((nelem + 255) / 256), dim3(256), 0, s, nelem, et.p, ep.p,
                       cn.p, mask.p);
  }
  if (nspc > 0) {
    sn.upload(spc_nodes_h, static_cast<size_t>(nspc), s);
    sm.upload(spc_masks_h, static_cast<size_t>(nspc), s);
    hipLaunchKernelGGL(k_dof_spc, dim3((nspc + 255) / 256), dim3(256), 0, s, nspc, sn.p, sm.p, spc.p);
  }
  cnt.resize(static_cast<size_t>(nnodes));
  base.resize(static_cast<size_t>(nnodes));
  hipLaunchKernelGGL(k_dof_count, dim3((nnodes + 255) / 256), dim3(256), 0, s, nnodes, order.p,
                     mask.p, spc.p, cnt.p);
  excl_scan(cnt.p, base.p, static_cast<size_t>(nnodes), s);
  const int nfree = read_last_plus(base.p, cnt.p, static_cast<size_t>(nnodes), s);
  eq.resize(static_cast<size_t>(nnodes) * 6);
  hipLaunchKernelGGL(k_dof_assign, dim3((nnodes + 255) / 256), dim3(256), 0, s, nnodes, order.p,
                     mask.p, spc.p, base.p, eq.p);
  VBH_HIP(hipMemcpyAsync(eq_out, eq.p, sizeof(int) * 6 * static_cast<size_t>(nnodes),
                         hipMemcpyDeviceToHost, s));
  VBH_HIP(hipStreamSynchronize(s));
  return nfree;
}

}  // namespace vbh

extern "C" {

int vbh_dof_map(vbh_ctx* h, int nnodes, const int* node_ids, int nelem, const int* elem_type,
                const int* elem_ptr, const int* conn, int nspc, const int* spc_nodes,
                const int* spc_masks, int* eq_out, int* n_free) {
  if (!h || !node_ids || !eq_out) {
    vbh::set_error("vbh_dof_map: null argument");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] {
    const int nf = c->dof_map(nnodes, node_ids, nelem, elem_type, elem_ptr, conn, nspc, spc_nodes,
                              spc_masks, eq_out);
    if (n_free) *n_free = nf;
  });
}

int vbh_assemble(vbh_ctx* h, int n, int nnodes, const double* xyz, const int* eq, int nelem,
                 const int* elem_type, const int* elem_ptr, const int* conn, const double* E,
                 const double* nu, const double* thickness, int solid_form, int shell_form) {
  if (!h || !xyz || !eq || !elem_type || !elem_ptr || !conn || !E || !nu) {
    vbh::set_error("vbh_assemble: null argument");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] {
    c->assemble(n, nnodes, xyz, eq, nelem, elem_type, elem_ptr, conn, E, nu, thickness, solid_form,
                shell_form);
  });
}

int vbh_element_sensitivity(vbh_ctx* h, int param, const double* u, const double* v,
                            double* out) {
  if (!h) {
    vbh::set_error("vbh_element_sensitivity: null ctx");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] { c->element_sensitivity(param, u, v, out, false); });
}

int vbh_element_sensitivity_device(vbh_ctx* h, int param, const double* du, const double* dv,
                                   double* dout) {
  if (!h) {
    vbh::set_error("vbh_element_sensitivity_device: null ctx");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] { c->element_sensitivity(param, du, dv, dout, true); });
}

int vbh_element_forces(vbh_ctx* h, const double* u, double* f) {
  if (!h) {
    vbh::set_error("vbh_element_forces: null ctx");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] { c->element_forces(u, f, false); });
}

int vbh_element_forces_device(vbh_ctx* h, const double* du, double* df) {
  if (!h) {
    vbh::set_error("vbh_element_forces_device: null ctx");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] { c->element_forces(du, df, true); });
}

int vbh_last_assembly_stats(const vbh_ctx* h, vbh_assembly_stats* out) {
  if (!h || !out) {
    vbh::set_error("vbh_last_assembly_stats: null argument");
    return VBH_ERR_ARG;
  }
  *out = reinterpret_cast<const vbh::Ctx*>(h)->asm_stats;
  return VBH_OK;
}

}  // extern "C"
