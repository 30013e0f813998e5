// vibestran_amd/csrc/vbh_mpc.hip -- multi-point constraint elimination.
//
// Host: the reference's master-slave elimination MpcHandler::build
// (src/core/mpc_handler.cpp:14-164) over an equation numbering: in every
// equation the term of largest |coefficient| (the first such) is dependent,
// u_dep = sum_j (-a_j / a_dep) u_j over the other terms; equations whose
// dependent DOF is constrained are skipped; dependent DOFs leave the numbering
// (later equations shift down) and terms on constrained or dependent DOFs drop
// out; a dependency cycle among dependent DOFs is an error. The result is the
// sparse map T (n_full x n_reduced): row i of a retained DOF is one 1.0 at its
// reduced index, row i of a dependent DOF holds its elimination terms.
//
// Device: K_red = T^T K_full T on the assembled lower CSR, kept in the full
// equation numbering with a unit diagonal on the dependent equations, so the
// operator's node blocks stay aligned (BSR / sliced ELL, mixed precision);
// the solver entry points map the reduced vectors in and out and vbh_get_csr
// returns the reduced CSR of the reference. Each stored entry
// (i, j, v) of K_full expands into the products of the T rows of i and j
// (c_a c_b v at (max(r_a, r_b), min(r_a, r_b)), doubled where both land on the
// same diagonal entry); the expanded list is sorted by (row, col) and every run
// of equal keys summed by one thread in sorted order, so the result is the
// same on every run. F_red = T^T F_full is a gather over the rows of T^T, and
// u_full = T u_red (MpcHandler::recover_dependent_dofs, mpc_handler.cpp:312-330)
// a gather over the rows of T in the reference's term order.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <unordered_map>
#include <unordered_set>

#include "vbh_ctx.hpp"

namespace vbh {

void MpcSystem::build(int nnodes, const int* eq, int n_full_, int nmpc, const int* ptr,
                      const int* node, const int* dof, const double* coeff) {
  *this = MpcSystem{};
  n_full = n_full_;
  if (nmpc <= 0) return;
  if (!eq || !ptr || !node || !dof || !coeff) throw_error(VBH_ERR_ARG, "vbh_set_mpc: null array");
  if (ptr[0] != 0) throw_error(VBH_ERR_ARG, "vbh_set_mpc: mpc_ptr[0] must be 0");
  for (int i = 0; i < 6 * nnodes; ++i)
    if (eq[i] < -1 || eq[i] >= n_full)
      throw_error(VBH_ERR_ARG, "vbh_set_mpc: eq entry out of range");
  struct Elim {
    int dep;
    std::vector<std::pair<int, double>> terms;
  };
  std::vector<Elim> elims;
  for (int k = 0; k < nmpc; ++k) {
    const int a = ptr[k], b = ptr[k + 1];
    if (b < a) throw_error(VBH_ERR_ARG, "vbh_set_mpc: mpc_ptr is not ascending");
    for (int t = a; t < b; ++t) {
      if (node[t] < 0 || node[t] >= nnodes)
        throw_error(VBH_ERR_ARG, "vbh_set_mpc: MPC term node index out of range");
      if (dof[t] < 1 || dof[t] > 6)
        throw_error(VBH_ERR_ARG, "vbh_set_mpc: MPC term component must be 1..6");
    }
    if (a == b) continue;
    int dep = a;
    double max_abs = 0.0;
    for (int t = a; t < b; ++t)
      if (std::abs(coeff[t]) > max_abs) {
        max_abs = std::abs(coeff[t]);
        dep = t;
      }
    if (max_abs == 0.0)
      throw_error(VBH_ERR_ARG, "vbh_set_mpc: MPC equation " + std::to_string(k) +
                                   " has no nonzero coefficient");
    const int dep_eq = eq[6 * node[dep] + dof[dep] - 1];
    if (dep_eq < 0) {
      std::fprintf(stderr,
                   "MPC equation %d: dependent DOF (node index %d, dof %d) is already "
                   "SPC-constrained; skipping MPC\n",
                   k, node[dep], dof[dep]);
      ++n_skipped;
      continue;
    }
    Elim e{dep_eq, {}};
    for (int t = a; t < b; ++t) {
      if (t == dep) continue;
      const int q = eq[6 * node[t] + dof[t] - 1];
      if (q < 0) continue;
      e.terms.emplace_back(q, -coeff[t] / coeff[dep]);
    }
    elims.push_back(std::move(e));
  }
  if (elims.empty()) return;

  // cycle check over the dependent-on-dependent graph (Kahn's algorithm)
  std::unordered_set<int> dep_set;
  for (const auto& e : elims) dep_set.insert(e.dep);
  {
    std::unordered_map<int, int> indeg;
    std::unordered_map<int, std::vector<int>> fwd;
    for (const auto& e : elims) indeg[e.dep] = 0;
    bool chain = false;
    for (const auto& e : elims)
      for (const auto& [q, c] : e.terms)
        if (dep_set.count(q)) {
          chain = true;
          fwd[q].push_back(e.dep);
          ++indeg[e.dep];
        }
    if (chain) {
      std::vector<int> queue;
      for (const auto& [q, d] : indeg)
        if (d == 0) queue.push_back(q);
      size_t done = 0;
      while (!queue.empty()) {
        const int cur = queue.back();
        queue.pop_back();
        ++done;
        for (int nx : fwd[cur])
          if (--indeg[nx] == 0) queue.push_back(nx);
      }
      if (done < indeg.size())
        throw_error(VBH_ERR_SOLVER,
                    "MPC dependency graph contains a cycle; check MPC equations for circular "
                    "references");
    }
  }

  // reduced numbering: dependent DOFs removed, the rest shifted down
  index_map.assign(static_cast<size_t>(n_full), -1);
  int removed = 0;
  for (int i = 0; i < n_full; ++i) {
    if (dep_set.count(i)) ++removed;
    else index_map[static_cast<size_t>(i)] = i - removed;
  }
  n_red = n_full - removed;
  red_to_full.reserve(static_cast<size_t>(n_red));
  for (int i = 0; i < n_full; ++i) {
    if (index_map[static_cast<size_t>(i)] >= 0) red_to_full.push_back(i);
    else deps.push_back(i);
  }
  // the last elimination of a dependent DOF defines its row
  std::vector<int> elim_of(static_cast<size_t>(n_full), -1);
  for (size_t k = 0; k < elims.size(); ++k) elim_of[static_cast<size_t>(elims[k].dep)] = static_cast<int>(k);
  t_ptr.assign(static_cast<size_t>(n_full) + 1, 0);
  for (int i = 0; i < n_full; ++i) {
    const int r = index_map[static_cast<size_t>(i)];
    if (r >= 0) {
      t_col.push_back(r);
      t_val.push_back(1.0);
    } else {
      for (const auto& [q, c] : elims[static_cast<size_t>(elim_of[static_cast<size_t>(i)])].terms) {
        const int rq = index_map[static_cast<size_t>(q)];
        if (rq < 0) continue;  // a term on another dependent DOF drops out
        t_col.push_back(rq);
        t_val.push_back(c);
      }
    }
    t_ptr[static_cast<size_t>(i) + 1] = static_cast<int>(t_col.size());
  }
  // T^T by rows: full rows in ascending order
  tt_ptr.assign(static_cast<size_t>(n_red) + 1, 0);
  for (int c : t_col) ++tt_ptr[static_cast<size_t>(c) + 1];
  for (int r = 0; r < n_red; ++r) tt_ptr[static_cast<size_t>(r) + 1] += tt_ptr[static_cast<size_t>(r)];
  tt_row.assign(t_col.size(), 0);
  tt_val.assign(t_col.size(), 0.0);
  std::vector<int> fill(tt_ptr.begin(), tt_ptr.end() - 1);
  for (int i = 0; i < n_full; ++i)
    for (int k = t_ptr[static_cast<size_t>(i)]; k < t_ptr[static_cast<size_t>(i) + 1]; ++k) {
      const int p = fill[static_cast<size_t>(t_col[static_cast<size_t>(k)])]++;
      tt_row[static_cast<size_t>(p)] = i;
      tt_val[static_cast<size_t>(p)] = t_val[static_cast<size_t>(k)];
    }
  n_elim = static_cast<int>(elims.size());
  active = true;
}

namespace {

// Products emitted by one lower entry (i, j) of K_full.
__device__ __forceinline__ long long mpc_products(int i, int j, const int* tptr) {
  const long long ni = tptr[i + 1] - tptr[i];
  if (i == j) return ni * (ni + 1) / 2;
  return ni * (tptr[j + 1] - tptr[j]);
}

__global__ void k_mpc_count(int nnz, const int* __restrict__ row, const int* __restrict__ col,
                            const int* __restrict__ tptr, long long* __restrict__ cnt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  cnt[e] = mpc_products(row[e], col[e], tptr);
}

// Expanded entries of lower entry e at off[e]: key (row << bits | col), value.
__global__ void k_mpc_emit(int nnz, const int* __restrict__ row, const int* __restrict__ col,
                           const double* __restrict__ v, const int* __restrict__ tptr,
                           const int* __restrict__ tcol, const double* __restrict__ tval,
                           const long long* __restrict__ off, int bits,
                           unsigned long long* __restrict__ keys, double* __restrict__ vals) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  const int i = row[e], j = col[e];
  const double x = v[e];
  long long o = off[e];
  for (int a = tptr[i]; a < tptr[i + 1]; ++a) {
    const int ra = tcol[a];
    const double ca = tval[a] * x;
    const int b_end = (i == j) ? a + 1 : tptr[j + 1];
    for (int b = tptr[j]; b < b_end; ++b) {
      const int rb = tcol[b];
      const bool twice = ra == rb && (i != j || a != b);
      const unsigned long long hi = static_cast<unsigned long long>(ra > rb ? ra : rb);
      const unsigned long long lo = static_cast<unsigned long long>(ra > rb ? rb : ra);
      keys[o] = (hi << bits) | lo;
      vals[o] = (twice ? 2.0 : 1.0) * ca * tval[b];
      ++o;
    }
  }
}

__global__ void k_mpc_heads(long long m, const unsigned long long* __restrict__ keys,
                            int* __restrict__ head) {
  const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (i >= m) return;
  head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// One thread per run of equal keys: the run's values summed in sorted order.
__global__ void k_mpc_runs(long long m, const unsigned long long* __restrict__ keys,
                           const double* __restrict__ vals, const int* __restrict__ head,
                           const int* __restrict__ pos, int bits, int* __restrict__ out_row,
                           int* __restrict__ out_col, double* __restrict__ out_v) {
  const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (i >= m || !head[i]) return;
  const unsigned long long k = keys[i];
  double s = vals[i];
  for (long long j = i + 1; j < m && keys[j] == k; ++j) s += vals[j];
  const int p = pos[i];
  out_row[p] = static_cast<int>(k >> bits);
  out_col[p] = static_cast<int>(k & ((1ull << bits) - 1ull));
  out_v[p] = s;
}

__global__ void k_mpc_row_ptr(int nnz, int n, const int* __restrict__ row_of, int* __restrict__ rp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nnz) return;
  const int prev = (i == 0) ? -1 : row_of[i - 1];
  const int cur = (i == nnz) ? n : row_of[i];
  for (int r = prev + 1; r <= cur; ++r) rp[r] = i;
}

// y[r] = sum_k val[k] x[idx[k]] over ptr[r]..ptr[r+1] (rows of T or T^T),
// each product rounded before it is added (no fused multiply-add), so u_full
// = T u_red is bit-identical to the reference's val += c * u_reduced[r] loop.
__global__ void k_mpc_gather(int n, const int* __restrict__ ptr, const int* __restrict__ idx,
                             const double* __restrict__ val, const double* __restrict__ x,
                             double* __restrict__ y) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int k = ptr[r]; k < ptr[r + 1]; ++k) s += val[k] * x[idx[k]];
  y[r] = s;
}

// Dependent rows of the padded operator: a unit diagonal, appended at off.
__global__ void k_mpc_dep_diag(int nd, const int* __restrict__ deps, int bits, long long off,
                               unsigned long long* __restrict__ keys, double* __restrict__ vals) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nd) return;
  const unsigned long long d = static_cast<unsigned long long>(deps[k]);
  keys[off + k] = (d << bits) | d;
  vals[off + k] = 1.0;
}

// dst[i] = map[i] >= 0 ? src[map[i]] : 0 for i < n, k columns of strides
// (ld_dst, ld_src).
__global__ void k_mpc_permute(int n, int k, const int* __restrict__ map, const double* __restrict__ src,
                              long long ld_src, double* __restrict__ dst, long long ld_dst) {
  const long long m = static_cast<long long>(n) * k;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < m;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int c = static_cast<int>(e / n), i = static_cast<int>(e % n);
    const int j = map[i];
    dst[c * ld_dst + i] = j >= 0 ? src[c * ld_src + j] : 0.0;
  }
}

}  // namespace

void MpcDev::upload(const MpcSystem& h, int n_mesh, hipStream_t s) {
  active = h.active;
  n_full = n_mesh;
  n_red = h.active ? h.n_red : n_mesh;
  if (!h.active) return;
  n_dep = static_cast<int>(h.deps.size());
  std::vector<int> colF(h.t_col.size());
  for (size_t k = 0; k < colF.size(); ++k) colF[k] = h.red_to_full[static_cast<size_t>(h.t_col[k])];
  t_ptr.upload(h.t_ptr.data(), h.t_ptr.size(), s);
  t_col.upload(h.t_col.data(), h.t_col.size(), s);
  t_colF.upload(colF.data(), colF.size(), s);
  deps.upload(h.deps.data(), h.deps.size(), s);
  t_val.upload(h.t_val.data(), h.t_val.size(), s);
  tt_ptr.upload(h.tt_ptr.data(), h.tt_ptr.size(), s);
  tt_row.upload(h.tt_row.data(), h.tt_row.size(), s);
  tt_val.upload(h.tt_val.data(), h.tt_val.size(), s);
  VBH_HIP(hipStreamSynchronize(s));
}

void MpcDev::expand(const double* u_red, double* u_full, hipStream_t s) const {
  if (n_full > 0)
    hipLaunchKernelGGL(k_mpc_gather, dim3((n_full + 255) / 256), dim3(256), 0, s, n_full,
                       t_ptr.p, t_col.p, t_val.p, u_red, u_full);
}

void MpcDev::reduce(const double* f_full, double* f_red, hipStream_t s) const {
  if (n_red > 0)
    hipLaunchKernelGGL(k_mpc_gather, dim3((n_red + 255) / 256), dim3(256), 0, s, n_red,
                       tt_ptr.p, tt_row.p, tt_val.p, f_full, f_red);
}

// K_red = T^T K_full T on the device lower CSR (l_rp/l_ci/l_v/l_row, n_full
// rows), replacing it. The result keeps the full numbering: retained
// equations carry K_red, dependent ones a unit diagonal and nothing else.
void Ctx::apply_mpc_to_lower() {
  const MpcDev& T = mesh.mpc;
  hipStream_t s = stream;
  const int nnz = l_nnz;
  DBuf<long long> cnt, off;
  cnt.resize(static_cast<size_t>(std::max(nnz, 1)));
  off.resize(static_cast<size_t>(std::max(nnz, 1)));
  if (nnz > 0)
    hipLaunchKernelGGL(k_mpc_count, dim3((nnz + 255) / 256), dim3(256), 0, s, nnz, l_row.p, l_ci.p,
                       T.t_ptr.p, cnt.p);
  long long m = 0;
  if (nnz > 0) {
    size_t tmp = 0;
    VBH_HIP(rocprim::exclusive_scan(nullptr, tmp, cnt.p, off.p, 0ll, static_cast<size_t>(nnz),
                                    rocprim::plus<long long>(), s));
    DBuf<char> t;
    t.resize(std::max<size_t>(tmp, 1));
    VBH_HIP(rocprim::exclusive_scan(t.p, tmp, cnt.p, off.p, 0ll, static_cast<size_t>(nnz),
                                    rocprim::plus<long long>(), s));
    long long a = 0, b = 0;
    VBH_HIP(hipMemcpyAsync(&a, off.p + nnz - 1, sizeof a, hipMemcpyDeviceToHost, s));
    VBH_HIP(hipMemcpyAsync(&b, cnt.p + nnz - 1, sizeof b, hipMemcpyDeviceToHost, s));
    VBH_HIP(hipStreamSynchronize(s));
    m = a + b;
  }
  const long long m_prod = m;
  m += T.n_dep;
  if (m >= (1ll << 31)) throw_error(VBH_ERR_SOLVER, "MPC transform: expanded operator too large");
  const int n_op = T.n_full;
  DBuf<unsigned long long> keys, keys_s;
  DBuf<double> vals, vals_s;
  const size_t mm = static_cast<size_t>(std::max<long long>(m, 1));
  keys.resize(mm);
  keys_s.resize(mm);
  vals.resize(mm);
  vals_s.resize(mm);
  int bits = 1;
  while ((1ll << bits) < static_cast<long long>(n_op) + 1) ++bits;
  if (nnz > 0)
    hipLaunchKernelGGL(k_mpc_emit, dim3((nnz + 255) / 256), dim3(256), 0, s, nnz, l_row.p, l_ci.p,
                       l_v.p, T.t_ptr.p, T.t_colF.p, T.t_val.p, off.p, bits, keys.p, vals.p);
  if (T.n_dep > 0)
    hipLaunchKernelGGL(k_mpc_dep_diag, dim3((T.n_dep + 255) / 256), dim3(256), 0, s, T.n_dep,
                       T.deps.p, bits, m_prod, keys.p, vals.p);
  if (m > 0) {
    // stable radix sort of (row << bits | col): equal keys keep their emission order
    size_t tmp = 0;
    VBH_HIP(rocprim::radix_sort_pairs(nullptr, tmp, keys.p, keys_s.p, vals.p, vals_s.p,
                                      static_cast<size_t>(m), 0u, 2u * static_cast<unsigned>(bits), s));
    DBuf<char> t;
    t.resize(std::max<size_t>(tmp, 1));
    VBH_HIP(rocprim::radix_sort_pairs(t.p, tmp, keys.p, keys_s.p, vals.p, vals_s.p,
                                      static_cast<size_t>(m), 0u, 2u * static_cast<unsigned>(bits), s));
  }
  keys.release();
  vals.release();
  DBuf<int> head, pos;
  head.resize(mm);
  pos.resize(mm);
  int nnz_red = 0;
  if (m > 0) {
    hipLaunchKernelGGL(k_mpc_heads, dim3(static_cast<unsigned>((m + 255) / 256)), dim3(256), 0, s, m,
                       keys_s.p, head.p);
    size_t tmp = 0;
    VBH_HIP(rocprim::exclusive_scan(nullptr, tmp, head.p, pos.p, 0, static_cast<size_t>(m),
                                    rocprim::plus<int>(), s));
    DBuf<char> t;
    t.resize(std::max<size_t>(tmp, 1));
    VBH_HIP(rocprim::exclusive_scan(t.p, tmp, head.p, pos.p, 0, static_cast<size_t>(m),
                                    rocprim::plus<int>(), s));
    int a = 0, b = 0;
    VBH_HIP(hipMemcpyAsync(&a, pos.p + m - 1, sizeof a, hipMemcpyDeviceToHost, s));
    VBH_HIP(hipMemcpyAsync(&b, head.p + m - 1, sizeof b, hipMemcpyDeviceToHost, s));
    VBH_HIP(hipStreamSynchronize(s));
    nnz_red = a + b;
  }
  l_ci.resize(static_cast<size_t>(std::max(nnz_red, 1)));
  l_v.resize(static_cast<size_t>(std::max(nnz_red, 1)));
  l_row.resize(static_cast<size_t>(std::max(nnz_red, 1)));
  if (m > 0)
    hipLaunchKernelGGL(k_mpc_runs, dim3(static_cast<unsigned>((m + 255) / 256)), dim3(256), 0, s, m,
                       keys_s.p, vals_s.p, head.p, pos.p, bits, l_row.p, l_ci.p, l_v.p);
  l_rp.resize(static_cast<size_t>(n_op) + 1);
  hipLaunchKernelGGL(k_mpc_row_ptr, dim3((nnz_red + 1 + 255) / 256), dim3(256), 0, s, nnz_red,
                     n_op, l_row.p, l_rp.p);
  op_api_host = mpc.index_map;
  op_api_of.upload(op_api_host.data(), op_api_host.size(), s);
  op_full_of.upload(mpc.red_to_full.data(), mpc.red_to_full.size(), s);
  VBH_HIP(hipStreamSynchronize(s));
  n = n_op;
  l_nnz = nnz_red;
  op_padded = true;
  n_api = T.n_red;
}

void Ctx::to_operator(const double* x, bool dev, double* dst, int k) {
  if (!op_padded) {
    VBH_HIP(hipMemcpyAsync(dst, x, sizeof(double) * n * k,
                           dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream));
    return;
  }
  DBuf<double> up;
  const double* src = x;
  if (!dev) {
    up.upload(x, static_cast<size_t>(n_api) * k, stream);
    src = up.p;
  }
  hipLaunchKernelGGL(k_mpc_permute, dim3(grid_for(static_cast<long long>(n) * k, 1024)), dim3(256), 0,
                     stream, n, k, op_api_of.p, src, static_cast<long long>(n_api), dst,
                     static_cast<long long>(n));
  if (!dev) VBH_HIP(hipStreamSynchronize(stream));
}

void Ctx::from_operator(const double* src, double* x, bool dev, int k) {
  if (!op_padded) {
    VBH_HIP(hipMemcpyAsync(x, src, sizeof(double) * n * k,
                           dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, stream));
    return;
  }
  DBuf<double> down;
  double* dst = x;
  if (!dev) {
    down.resize(static_cast<size_t>(n_api) * k);
    dst = down.p;
  }
  hipLaunchKernelGGL(k_mpc_permute, dim3(grid_for(static_cast<long long>(n_api) * k, 1024)),
                     dim3(256), 0, stream, n_api, k, op_full_of.p, src, static_cast<long long>(n),
                     dst, static_cast<long long>(n_api));
  if (!dev) {
    VBH_HIP(hipMemcpyAsync(x, down.p, sizeof(double) * n_api * k, hipMemcpyDeviceToHost, stream));
    VBH_HIP(hipStreamSynchronize(stream));
  }
}

// u in the operator's numbering (host or device) as a device vector in the
// mesh numbering: uploaded when on the host, expanded through T under MPCs.
const double* Ctx::mesh_solution(const double* u, bool dev, DBuf<double>& up,
                                 DBuf<double>& full) {
  const double* p = u;
  if (!dev) {
    up.upload(u, static_cast<size_t>(mesh.mpc.n_red), stream);
    p = up.p;
  }
  if (!mesh.mpc.active) return p;
  full.resize(static_cast<size_t>(std::max(mesh.n, 1)));
  mesh.mpc.expand(p, full.p, stream);
  return full.p;
}

}  // namespace vbh

extern "C" {

int vbh_set_mpc(vbh_ctx* h, int nnodes, const int* eq, int n_full, int nmpc, const int* mpc_ptr,
                const int* term_node, const int* term_dof, const double* term_coeff,
                int* n_reduced) {
  if (!h) {
    vbh::set_error("vbh_set_mpc: null ctx");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] {
    if (nnodes < 0 || n_full < 0 || nmpc < 0) vbh::throw_error(VBH_ERR_ARG, "vbh_set_mpc: negative size");
    vbh::MpcSystem sys;
    sys.build(nnodes, eq, n_full, nmpc, mpc_ptr, term_node, term_dof, term_coeff);
    c->mpc = std::move(sys);
    if (n_reduced) *n_reduced = c->mpc.active ? c->mpc.n_red : n_full;
  });
}

int vbh_mpc_info(const vbh_ctx* h, vbh_mpc_stats* out) {
  if (!h || !out) {
    vbh::set_error("vbh_mpc_info: null argument");
    return VBH_ERR_ARG;
  }
  const auto* c = reinterpret_cast<const vbh::Ctx*>(h);
  out->n_full = c->mpc.n_full;
  out->n_reduced = c->mpc.active ? c->mpc.n_red : c->mpc.n_full;
  out->n_eliminations = c->mpc.n_elim;
  out->n_skipped = c->mpc.n_skipped;
  out->t_nnz = static_cast<int>(c->mpc.t_col.size());
  return VBH_OK;
}

int vbh_mpc_transform(const vbh_ctx* h, int* t_ptr, int* t_col, double* t_val) {
  if (!h) {
    vbh::set_error("vbh_mpc_transform: null ctx");
    return VBH_ERR_ARG;
  }
  const auto* c = reinterpret_cast<const vbh::Ctx*>(h);
  if (!c->mpc.active) {
    vbh::set_error("vbh_mpc_transform: no MPC elimination is set");
    return VBH_ERR_STATE;
  }
  if (t_ptr) std::copy(c->mpc.t_ptr.begin(), c->mpc.t_ptr.end(), t_ptr);
  if (t_col) std::copy(c->mpc.t_col.begin(), c->mpc.t_col.end(), t_col);
  if (t_val) std::copy(c->mpc.t_val.begin(), c->mpc.t_val.end(), t_val);
  return VBH_OK;
}

int vbh_mpc_expand(vbh_ctx* h, const double* u_red, double* u_full) {
  if (!h) {
    vbh::set_error("vbh_mpc_expand: null ctx");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] {
    if (!c->mesh.valid) vbh::throw_error(VBH_ERR_STATE, "vbh_mpc_expand: no assembled mesh in this context");
    if (!u_red || !u_full) vbh::throw_error(VBH_ERR_ARG, "vbh_mpc_expand: null vector");
    c->select_stream_device();
    vbh::DBuf<double> up, full;
    const double* p = c->mesh_solution(u_red, false, up, full);
    VBH_HIP(hipMemcpyAsync(u_full, p, sizeof(double) * c->mesh.n, hipMemcpyDeviceToHost, c->stream));
    VBH_HIP(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"
