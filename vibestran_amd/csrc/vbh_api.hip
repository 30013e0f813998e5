// vibestran_amd/csrc/vbh_api.hip -- C ABI of libvbh: context, operator upload,
// Jacobi-PCG drivers (FP64 and FP32 + FP64 refresh), diagnostics.
//
// Drop-in for the reference's SolverBackend plugin (reference
// include/solver/solver_backend.hpp:20-51); error texts follow the reference's
// SolverError messages (cuda_pcg_backend_impl.hpp:1964-1970, 2076-2114).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <utility>
#include <vector>

#include "vbh_ctx.hpp"

namespace vbh {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

void throw_error(int code, const std::string& msg) { throw Error{code, msg}; }

std::string fmt_g(double v) {
  char b[64];
  std::snprintf(b, sizeof b, "%g", v);
  return b;
}

// ---------------------------------------------------------------------------
// Operator construction on the host from the reference's CsrData layout.
// ---------------------------------------------------------------------------
namespace {

struct HostFull {
  int n = 0;
  std::vector<int> rp, ci;
  std::vector<double> v;
};

// Lower (or full) CSR -> full symmetric CSR with sorted rows; same result as
// CsrData::expanded_symmetric (sparse_matrix.cpp:183-235).
HostFull expand_full(int n, int nnz, const int* rp, const int* ci, const double* v, bool lower) {
  HostFull f;
  f.n = n;
  if (!lower) {
    f.rp.assign(rp, rp + n + 1);
    f.ci.assign(ci, ci + nnz);
    f.v.assign(v, v + nnz);
    return f;
  }
  f.rp.assign(static_cast<size_t>(n) + 1, 0);
  for (int r = 0; r < n; ++r)
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
      ++f.rp[static_cast<size_t>(r) + 1];
      if (ci[k] != r) ++f.rp[static_cast<size_t>(ci[k]) + 1];
    }
  std::partial_sum(f.rp.begin(), f.rp.end(), f.rp.begin());
  const size_t tot = static_cast<size_t>(f.rp[static_cast<size_t>(n)]);
  f.ci.resize(tot);
  f.v.resize(tot);
  std::vector<int> cur(f.rp.begin(), f.rp.begin() + n);
  for (int r = 0; r < n; ++r)
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
      const int c = ci[k];
      int o = cur[static_cast<size_t>(r)]++;
      f.ci[static_cast<size_t>(o)] = c;
      f.v[static_cast<size_t>(o)] = v[k];
      if (c != r) {
        o = cur[static_cast<size_t>(c)]++;
        f.ci[static_cast<size_t>(o)] = r;
        f.v[static_cast<size_t>(o)] = v[k];
      }
    }
  return f;
}

// Rows of f sorted by column with repeated entries summed in storage order
// (callers may pass unsorted rows or repeated entries; the reference's
// CsrData, built by build_csr, has neither). Rows already strictly
// increasing are copied as they are.
void canonical_rows(HostFull& f) {
  bool clean = true;
  for (int r = 0; r < f.n && clean; ++r)
    for (int k = f.rp[static_cast<size_t>(r)] + 1; k < f.rp[static_cast<size_t>(r) + 1]; ++k)
      if (f.ci[static_cast<size_t>(k)] <= f.ci[static_cast<size_t>(k) - 1]) {
        clean = false;
        break;
      }
  if (clean) return;
  std::vector<int> rp(static_cast<size_t>(f.n) + 1, 0), ci;
  std::vector<double> v;
  ci.reserve(f.ci.size());
  v.reserve(f.v.size());
  std::vector<std::pair<int, int>> row;  // (column, position)
  for (int r = 0; r < f.n; ++r) {
    row.clear();
    for (int k = f.rp[static_cast<size_t>(r)]; k < f.rp[static_cast<size_t>(r) + 1]; ++k)
      row.emplace_back(f.ci[static_cast<size_t>(k)], k);
    std::stable_sort(row.begin(), row.end(),
                     [](const std::pair<int, int>& a, const std::pair<int, int>& b) { return a.first < b.first; });
    for (size_t q = 0; q < row.size(); ++q) {
      const double x = f.v[static_cast<size_t>(row[q].second)];
      if (q > 0 && row[q].first == row[q - 1].first) {
        v.back() += x;
      } else {
        ci.push_back(row[q].first);
        v.push_back(x);
      }
    }
    rp[static_cast<size_t>(r) + 1] = static_cast<int>(ci.size());
  }
  f.rp.swap(rp);
  f.ci.swap(ci);
  f.v.swap(v);
}

struct HostBsr {
  int B = 0, nbr = 0;
  std::vector<int> brp, bci;
  std::vector<double> v;  // entry-major per block row
};

HostBsr to_bsr(const HostFull& f, int B) {
  HostBsr b;
  b.B = B;
  b.nbr = (f.n + B - 1) / B;
  b.brp.assign(static_cast<size_t>(b.nbr) + 1, 0);
  std::vector<int> cols;
  std::vector<std::vector<int>> rows_cols(1);
  // pass 1: block pattern
  for (int I = 0; I < b.nbr; ++I) {
    cols.clear();
    for (int r = B * I; r < std::min(f.n, B * I + B); ++r)
      for (int k = f.rp[static_cast<size_t>(r)]; k < f.rp[static_cast<size_t>(r) + 1]; ++k)
        cols.push_back(f.ci[static_cast<size_t>(k)] / B);
    cols.push_back(I);
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    b.bci.insert(b.bci.end(), cols.begin(), cols.end());
    b.brp[static_cast<size_t>(I) + 1] = static_cast<int>(b.bci.size());
  }
  b.v.assign(b.bci.size() * static_cast<size_t>(B * B), 0.0);
  // pass 2: values
  for (int I = 0; I < b.nbr; ++I) {
    const int o = b.brp[static_cast<size_t>(I)], nb = b.brp[static_cast<size_t>(I) + 1] - o;
    const int* seg = b.bci.data() + o;
    for (int r = B * I; r < B * I + B; ++r) {
      if (r >= f.n) {  // padding row: identity diagonal
        const int k = static_cast<int>(std::lower_bound(seg, seg + nb, I) - seg);
        const int e = (r % B) * B + (r % B);
        b.v[static_cast<size_t>(B * B) * o + static_cast<size_t>(e) * nb + k] = 1.0;
        continue;
      }
      for (int q = f.rp[static_cast<size_t>(r)]; q < f.rp[static_cast<size_t>(r) + 1]; ++q) {
        const int c = f.ci[static_cast<size_t>(q)];
        const int k = static_cast<int>(std::lower_bound(seg, seg + nb, c / B) - seg);
        const int e = (r % B) * B + (c % B);
        b.v[static_cast<size_t>(B * B) * o + static_cast<size_t>(e) * nb + k] += f.v[static_cast<size_t>(q)];
      }
    }
  }
  return b;
}

}  // namespace

void Ctx::select_stream_device() { VBH_HIP(hipSetDevice(device)); }

void Ctx::upload_operator(int n_, int nnz, const int* rp, const int* ci, const double* v,
                          bool lower) {
  select_stream_device();
  if (n_ <= 0) throw_error(VBH_ERR_SOLVER, "HIP PCG: stiffness matrix is empty -- no free DOFs");
  if (!rp || !ci || (!v && nnz > 0)) throw_error(VBH_ERR_ARG, "vbh_set_csr: null array");
  if (rp[0] != 0 || rp[n_] != nnz) throw_error(VBH_ERR_ARG, "vbh_set_csr: row_ptr inconsistent with nnz");
  for (int i = 0; i < n_; ++i)
    if (rp[i + 1] < rp[i]) throw_error(VBH_ERR_ARG, "vbh_set_csr: row_ptr decreases");
  if ((lower ? 2ll * nnz - n_ : static_cast<long long>(nnz)) >= 0x7fffffffll)
    throw_error(VBH_ERR_ARG, "vbh_set_csr: the full symmetric operator has 2^31 or more entries "
                             "(32-bit indices)");
  // Sorted lower storage (the reference's CsrData) is converted on the device.
  if (lower && format_req != 0 && set_lower_device(n_, nnz, rp, ci, v)) return;
  // Jacobi diagonal with the reference's checks (build_jacobi_diagonal).
  std::vector<double> diag(static_cast<size_t>(n_), 0.0);
  for (int i = 0; i < n_; ++i) {
    bool found = false;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      if (ci[k] < 0 || ci[k] >= n_) throw_error(VBH_ERR_ARG, "vbh_set_csr: column index out of range");
      if (lower && ci[k] > i) throw_error(VBH_ERR_ARG, "vbh_set_csr: upper entry in lower storage");
      if (ci[k] == i && !found) {
        if (v[k] == 0.0)
          throw_error(VBH_ERR_SOLVER, "HIP PCG: zero diagonal at row " + std::to_string(i) +
                                          " -- matrix is singular. Check boundary conditions.");
        diag[static_cast<size_t>(i)] = v[k];
        found = true;
      }
    }
    if (!found)
      throw_error(VBH_ERR_SOLVER, "HIP PCG: missing diagonal at row " + std::to_string(i) +
                                      " -- matrix is singular. Check boundary conditions.");
  }
  n = n_;
  // Keep the caller's lower CSR (reference CsrData) for vbh_get_csr.
  lower_on_host = true;
  if (lower) {
    lower_rp.assign(rp, rp + n_ + 1);
    lower_ci.assign(ci, ci + nnz);
    lower_v.assign(v, v + nnz);
  } else {
    lower_rp.assign(static_cast<size_t>(n_) + 1, 0);
    lower_ci.clear();
    lower_v.clear();
    for (int r = 0; r < n_; ++r) {
      for (int k = rp[r]; k < rp[r + 1]; ++k)
        if (ci[k] <= r) {
          lower_ci.push_back(ci[k]);
          lower_v.push_back(v[k]);
        }
      lower_rp[static_cast<size_t>(r) + 1] = static_cast<int>(lower_ci.size());
    }
  }
  HostFull full = expand_full(n_, nnz, rp, ci, v, lower);
  canonical_rows(full);
  // Format choice: BSR when the block fill stays low.
  int fmt = format_req;
  HostBsr bsr;
  if (fmt < 0) {
    fmt = 0;
    // 6x6 node blocks when they carry no more padding than 3x3 blocks would
    // (shell meshes: structural zeros make both about 1.2), else 3x3.
    const double nv = static_cast<double>(full.v.size());
    HostBsr b6, b3;
    double f6 = 1e300, f3 = 1e300;
    if (n_ >= 24) {
      b6 = to_bsr(full, 6);
      f6 = static_cast<double>(b6.v.size()) / nv;
    }
    if (f6 > 1.12 && n_ >= 12) {
      b3 = to_bsr(full, 3);
      f3 = static_cast<double>(b3.v.size()) / nv;
    }
    if (f6 <= 1.12 || (f6 <= 1.3 && f6 <= 1.05 * f3)) {
      bsr = std::move(b6);
      fmt = 6;
    } else if (f3 <= 1.25) {
      bsr = std::move(b3);
      fmt = 3;
    }
  } else if (fmt == 3 || fmt == 6) {
    bsr = to_bsr(full, fmt);
  } else {
    fmt = 0;
  }
  format = fmt;
  n_pad = (fmt == 0) ? n_ : bsr.nbr * fmt;
  hipStream_t s = stream;
  // scalar full CSR (always kept: generic path and FP64 reference operator)
  csr_rp.upload(full.rp.data(), full.rp.size(), s);
  csr_ci.upload(full.ci.data(), full.ci.size(), s);
  csr_v64.upload(full.v.data(), full.v.size(), s);
  {
    std::vector<float> f32(full.v.begin(), full.v.end());
    csr_v32.upload(f32.data(), f32.size(), s);
    VBH_HIP(hipStreamSynchronize(s));
  }
  csr.n = n_;
  csr.nnz = static_cast<int>(full.v.size());
  csr.rp = csr_rp.p;
  csr.ci = csr_ci.p;
  csr.v64 = csr_v64.p;
  csr.v32 = csr_v32.p;
  if (fmt != 0) {
    bsr_rp.upload(bsr.brp.data(), bsr.brp.size(), s);
    bsr_ci.upload(bsr.bci.data(), bsr.bci.size(), s);
    bsr_v64.upload(bsr.v.data(), bsr.v.size(), s);
    std::vector<float> f32(bsr.v.begin(), bsr.v.end());
    bsr_v32.upload(f32.data(), f32.size(), s);
    VBH_HIP(hipStreamSynchronize(s));
    bsrd.B = fmt;
    bsrd.nbr = bsr.nbr;
    bsrd.nnzb = static_cast<int>(bsr.bci.size());
    bsrd.brp = bsr_rp.p;
    bsrd.bci = bsr_ci.p;
    bsrd.v64 = bsr_v64.p;
    bsrd.v32 = bsr_v32.p;
    build_sell(*this);
  } else {
    bsr_rp.release();
    bsr_ci.release();
    bsr_v64.release();
    bsr_v32.release();
    bsrd = BsrDev{};
  }
  diag.resize(static_cast<size_t>(n_pad), 1.0);
  diag64.upload(diag.data(), diag.size(), s);
  {
    std::vector<float> d32(diag.begin(), diag.end());
    diag32.upload(d32.data(), d32.size(), s);
    VBH_HIP(hipStreamSynchronize(s));
  }
  l_nnz = static_cast<int>(lower_ci.size());
  operator_version += 1;
  have_matrix = true;
}

// ---------------------------------------------------------------------------
// Solve driver
// ---------------------------------------------------------------------------
void Ctx::ensure_work(int nvec_len) {
  const size_t N = static_cast<size_t>(nvec_len);
  if (work_len == N && work_version == operator_version) return;
  w_F.resize(N);
  w_u.resize(N);
  w_r64.resize(N);
  w_r.resize(N);
  w_p.resize(N);
  w_Ap.resize(N);
  w_r32.resize(N);
  w_p32.resize(N);
  w_Ap32.resize(N);
  if (!part_a.p) {
    part_a.resize(static_cast<size_t>(kMaxRed) * kMaxGrid);
    part_b.resize(2 * static_cast<size_t>(kPartB));
  }
  if (!st_dev.p) st_dev.resize(1);
  if (!st_host)
    VBH_HIP(hipHostMalloc(reinterpret_cast<void**>(&st_host), 2 * sizeof(PcgScalars), 0));
  // zero the padding tails once (F beyond n stays zero)
  launch_fill_f64(static_cast<long long>(N), w_F.p, 0.0, stream);
  const int rows = (format == 0) ? n : bsrd.nslices;
  red_n_a = grid_for(rows, rows_per_block(format));
  red_n_b = vector_grid(static_cast<long long>(N), n_cu);
  drop_graphs();
  work_len = N;
  work_version = operator_version;
}

void Ctx::drop_graphs() {
  for (auto& g : graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
  }
}

PcgParams Ctx::params() const {
  PcgParams p;
  p.tol = tol;
  p.max_iters = max_iters;
  p.refresh_delta = delta();
  return p;
}

CgRed Ctx::red() const {
  CgRed r;
  r.partA = part_a.p;
  r.partB = part_b.p;
  r.nA = red_n_a;
  r.nB = red_n_b;
  return r;
}

CgBufs<double> Ctx::bufs64() {
  return CgBufs<double>{w_F.p, diag64.p, w_r.p, w_p.p, w_Ap.p, w_u.p, nullptr,
                        format != 0 ? bsrd.xcd[0] : 0};
}

CgBufs<float> Ctx::bufs32() {
  return CgBufs<float>{w_F.p, diag32.p, w_r32.p, w_p32.p, w_Ap32.p, w_u.p, w_r64.p,
                       format != 0 ? bsrd.xcd[1] : 0};
}

void Ctx::enqueue_iterations(int precision_mode, int count) {
  if (count % 2 != 0) throw_error(VBH_ERR_ARG, "HIP PCG: iteration launch sequences have an even length");
  const int N = static_cast<int>(work_len);
  const PcgParams prm = params();
  const CgRed rd = red();
  for (int i = 0; i < count; ++i) {
    if (precision_mode == VBH_PREC_FP64) {
      const CgBufs<double> b = bufs64();
      launch_cg_spmv<double>(format, csr, bsrd, b, st_dev.p, rd, stream, 0, i);
      launch_cg_vector<double>(N, b, st_dev.p, rd, prm, 0, i, stream);
    } else {
      const CgBufs<float> b = bufs32();
      launch_cg_spmv<float>(format, csr, bsrd, b, st_dev.p, rd, stream, spmv_acc64, i);
      launch_cg_vector<float>(N, b, st_dev.p, rd, prm, 1, i, stream);
    }
  }
}

hipGraphExec_t Ctx::iteration_graph(int precision_mode) {
  GraphSlot& g = graphs[precision_mode];
  if (g.exec && g.iters == iters_per_graph && g.tol == tol && g.max_iters == max_iters)
    return g.exec;
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  g.exec = nullptr;
  hipGraph_t graph = nullptr;
  VBH_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  enqueue_iterations(precision_mode, iters_per_graph);
  VBH_HIP(hipStreamEndCapture(stream, &graph));
  VBH_HIP(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
  (void)hipGraphDestroy(graph);
  g.iters = iters_per_graph;
  g.tol = tol;
  g.max_iters = max_iters;
  return g.exec;
}

// Launches batches of iterations until the device state reports done. The
// state of batch k is copied out behind it and read while batch k + 1 is
// already queued, so the device never waits for the host; the batch queued
// behind the last one finds the state stopped and every kernel of it exits at
// entry. A batch that ends neither stopped nor short of its iterations always
// advances the state, so the cap is never reached by a healthy solve
// (iterations plus refresh steps are at most twice the iterations).
PcgScalars Ctx::drive(const PcgScalars* st, const std::function<void()>& batch) {
  const long long cap = 2ll * (max_iters / iters_per_graph + 2) + 64;
  long long launches = 0;
  auto issue = [&](int slot) {
    if (++launches > cap)
      throw_error(VBH_ERR_HIP, "HIP PCG: device iteration state stopped advancing");
    batch();
    VBH_HIP(hipMemcpyAsync(st_host + slot, st, sizeof(PcgScalars), hipMemcpyDeviceToHost, stream));
    VBH_HIP(hipEventRecord(ev_poll[slot], stream));
  };
  issue(0);
  int cur = 0;
  for (;;) {
    issue(cur ^ 1);
    VBH_HIP(hipEventSynchronize(ev_poll[cur]));
    if (st_host[cur].done) break;
    cur ^= 1;
  }
  VBH_HIP(hipEventSynchronize(ev_poll[cur ^ 1]));
  return st_host[cur ^ 1];
}

// Runs one solve on device-resident F (w_F) into w_u. Returns the scalars.
PcgScalars Ctx::run_solve(int precision_mode) {
  const int N = static_cast<int>(work_len);
  const CgRed rd = red();
  VBH_HIP(hipEventRecord(ev_start, stream));
  if (precision_mode == VBH_PREC_FP64)
    launch_cg_init<double>(N, bufs64(), st_dev.p, rd, stream);
  else
    launch_cg_init<float>(N, bufs32(), st_dev.p, rd, stream);
  const char* ge_env = std::getenv("VBH_USE_GRAPH");
  const bool use_graph = !(ge_env && ge_env[0] == '0');
  hipGraphExec_t ge = use_graph ? iteration_graph(precision_mode) : nullptr;
  const PcgScalars s = drive(st_dev.p, [&] {
    if (use_graph) VBH_HIP(hipGraphLaunch(ge, stream));
    else enqueue_iterations(precision_mode, iters_per_graph);
  });
  VBH_HIP(hipEventRecord(ev_stop, stream));
  VBH_HIP(hipEventSynchronize(ev_stop));
  float ms = 0.0f;
  VBH_HIP(hipEventElapsedTime(&ms, ev_start, ev_stop));
  stats.iterations = s.iter_done;
  stats.refinements = s.refreshes;
  stats.rel_res = (precision_mode == VBH_PREC_MIXED && (s.flags & (kFlagConverged | kFlagStall)))
                      ? s.rel_true
                      : s.rel;
  if (s.flags & kFlagZeroRhs) stats.rel_res = 0.0;
  stats.solve_ms = ms;
  stats.format = format;
  return s;
}

// Maps the device flags onto the reference's outcomes.
void Ctx::check_outcome(const PcgScalars& s, int precision_mode) {
  if (s.flags & kFlagRz0)
    throw_error(VBH_ERR_SOLVER, "HIP PCG: non-positive initial r*z=" + fmt_g(s.rz[0]) +
                                    " -- preconditioner is not SPD or matrix is singular");
  if (s.flags & (kFlagBreakdown | kFlagMaxIter)) {
    // Breakdown is reported as non-convergence (impl:1934-1942, 1964-1970).
    const double rel = (s.flags & kFlagBreakdown) ? 1e300 : s.rel;
    last_iters = (s.flags & kFlagBreakdown) ? max_iters : s.iter_done;
    last_rel = rel;
    throw_error(VBH_ERR_SOLVER,
                "HIP PCG: did not converge after " + std::to_string(max_iters) +
                    " iterations (relative residual " + fmt_g(rel) + " > tolerance " + fmt_g(tol) +
                    "). Consider increasing max iterations or using a direct solver.");
  }
  (void)precision_mode;
  last_iters = s.iter_done;
  last_rel = stats.rel_res;
}

}  // namespace vbh

// ===========================================================================
// extern "C"
// ===========================================================================
using vbh::Ctx;
using vbh::guarded;
using vbh::DBuf;

extern "C" {

const char* vbh_last_error(void) { return vbh::g_last_error.c_str(); }

int vbh_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

int vbh_ctx_create(int device, double tol, int max_iters, vbh_ctx** out) {
  if (!out) {
    vbh::set_error("vbh_ctx_create: null out");
    return VBH_ERR_ARG;
  }
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    vbh::set_error("no HIP device available");
    return VBH_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= count) {
    vbh::set_error("vbh_ctx_create: device index out of range");
    return VBH_ERR_NO_DEVICE;
  }
  return guarded([&] {
    auto c = std::make_unique<Ctx>();
    c->device = device;
    VBH_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    VBH_HIP(hipGetDeviceProperties(&prop, device));
    c->name = prop.name;
    if (c->name.empty()) {
      std::string arch = prop.gcnArchName;
      c->name = "AMD Instinct " + arch.substr(0, arch.find(':')) + " (" +
                std::to_string(prop.multiProcessorCount) + " CUs)";
    }
    c->arch = prop.gcnArchName;
    c->n_cu = prop.multiProcessorCount;
    c->tol = tol > 0.0 ? tol : 1e-8;
    c->max_iters = max_iters > 0 ? max_iters : 10000;
    if (const char* e = std::getenv("VBH_REFRESH_DELTA")) c->refresh_delta = std::atof(e);
    if (const char* e = std::getenv("VBH_SPMV_ACC64")) c->spmv_acc64 = std::atoi(e);
    if (const char* e = std::getenv("VBH_ROWSUM")) c->rowsum_fp32 = std::atoi(e);
    VBH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    VBH_HIP(hipEventCreate(&c->ev_start));
    VBH_HIP(hipEventCreate(&c->ev_stop));
    for (hipEvent_t& e : c->ev_poll) VBH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // set-up for the allocation-free solve: the pinned state word slots and
    // the device address of this device's solver scratch array
    VBH_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->st_host), 2 * sizeof(vbh::PcgScalars), 0));
    vbh::scratch_init(device);
    *out = reinterpret_cast<vbh_ctx*>(c.release());
  });
}

void vbh_ctx_destroy(vbh_ctx* h) {
  if (!h) return;
  delete reinterpret_cast<Ctx*>(h);
}

const char* vbh_device_name(const vbh_ctx* h) {
  return h ? reinterpret_cast<const Ctx*>(h)->name.c_str() : "";
}

int vbh_set_precision(vbh_ctx* h, int precision) {
  if (!h || (precision != VBH_PREC_FP64 && precision != VBH_PREC_MIXED)) {
    vbh::set_error("vbh_set_precision: bad argument");
    return VBH_ERR_ARG;
  }
  reinterpret_cast<Ctx*>(h)->precision = precision;
  return VBH_OK;
}

int vbh_set_tolerance(vbh_ctx* h, double tol, int max_iters) {
  if (!h) {
    vbh::set_error("vbh_set_tolerance: null ctx");
    return VBH_ERR_ARG;
  }
  Ctx* c = reinterpret_cast<Ctx*>(h);
  c->tol = tol > 0.0 ? tol : 1e-8;
  c->max_iters = max_iters > 0 ? max_iters : 10000;
  return VBH_OK;
}

int vbh_set_format(vbh_ctx* h, int format) {
  if (!h || !(format == -1 || format == 0 || format == 3 || format == 6)) {
    vbh::set_error("vbh_set_format: format must be -1, 0, 3 or 6");
    return VBH_ERR_ARG;
  }
  reinterpret_cast<Ctx*>(h)->format_req = format;
  return VBH_OK;
}

int vbh_set_csr(vbh_ctx* h, int n, int nnz, const int* row_ptr, const int* col_ind,
                const double* values, int lower_only) {
  if (!h) {
    vbh::set_error("vbh_set_csr: null ctx");
    return VBH_ERR_ARG;
  }
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded([&] {
    c->upload_operator(n, nnz, row_ptr, col_ind, values, lower_only != 0);
    c->op_padded = false;
    c->n_api = c->n;
  });
}

int vbh_get_csr_size(const vbh_ctx* h, int* n, int* nnz) {
  if (!h) {
    vbh::set_error("vbh_get_csr_size: null ctx");
    return VBH_ERR_ARG;
  }
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c->have_matrix) {
    vbh::set_error("no matrix set");
    return VBH_ERR_STATE;
  }
  if (n) *n = c->n_api;
  if (nnz) *nnz = c->op_padded ? c->l_nnz - (c->n - c->n_api) : c->l_nnz;
  return VBH_OK;
}

int vbh_get_csr(const vbh_ctx* h, int* row_ptr, int* col_ind, double* values) {
  if (!h) {
    vbh::set_error("vbh_get_csr: null ctx");
    return VBH_ERR_ARG;
  }
  Ctx* c = const_cast<Ctx*>(reinterpret_cast<const Ctx*>(h));
  if (!c->have_matrix) {
    vbh::set_error("no matrix set");
    return VBH_ERR_STATE;
  }
  return guarded([&] {
    c->fetch_lower_to_host();
    if (!c->op_padded) {
      if (row_ptr) std::copy(c->lower_rp.begin(), c->lower_rp.end(), row_ptr);
      if (col_ind) std::copy(c->lower_ci.begin(), c->lower_ci.end(), col_ind);
      if (values) std::copy(c->lower_v.begin(), c->lower_v.end(), values);
      return;
    }
    // the reduced CSR: dependent rows (a unit diagonal) dropped, the other
    // rows and columns renumbered (no retained entry names a dependent column)
    const std::vector<int>& api_of = c->op_api_host;
    int r = 0, k = 0;
    if (row_ptr) row_ptr[0] = 0;
    for (int i = 0; i < c->n; ++i) {
      if (api_of[static_cast<size_t>(i)] < 0) continue;
      for (int e = c->lower_rp[static_cast<size_t>(i)]; e < c->lower_rp[static_cast<size_t>(i) + 1]; ++e, ++k) {
        if (col_ind) col_ind[k] = api_of[static_cast<size_t>(c->lower_ci[static_cast<size_t>(e)])];
        if (values) values[k] = c->lower_v[static_cast<size_t>(e)];
      }
      if (row_ptr) row_ptr[++r] = k;
    }
  });
}

static int solve_impl(Ctx* c, const double* F, double* u, bool device_ptrs, int* iters,
                      double* rel_res) {
  return guarded([&] {
    if (!c->have_matrix) vbh::throw_error(VBH_ERR_STATE, "vbh_solve: no matrix set");
    if (!F || !u) vbh::throw_error(VBH_ERR_ARG, "vbh_solve: null vector");
    c->select_stream_device();
    c->ensure_work(c->n_pad);
    c->to_operator(F, device_ptrs, c->w_F.p);
    vbh::PcgScalars s = c->run_solve(c->precision);
    c->from_operator(c->w_u.p, u, device_ptrs);
    VBH_HIP(hipStreamSynchronize(c->stream));
    if (iters) *iters = s.iter_done;
    if (rel_res) *rel_res = c->stats.rel_res;
    c->check_outcome(s, c->precision);
  });
}

int vbh_solve(vbh_ctx* h, const double* F, double* u, int* iters, double* rel_res) {
  if (!h) {
    vbh::set_error("vbh_solve: null ctx");
    return VBH_ERR_ARG;
  }
  return solve_impl(reinterpret_cast<Ctx*>(h), F, u, false, iters, rel_res);
}

int vbh_solve_device(vbh_ctx* h, const double* dF, double* du, int* iters, double* rel_res) {
  if (!h) {
    vbh::set_error("vbh_solve_device: null ctx");
    return VBH_ERR_ARG;
  }
  return solve_impl(reinterpret_cast<Ctx*>(h), dF, du, true, iters, rel_res);
}

int vbh_last_stats(const vbh_ctx* h, vbh_solve_stats* out) {
  if (!h || !out) {
    vbh::set_error("vbh_last_stats: null argument");
    return VBH_ERR_ARG;
  }
  *out = reinterpret_cast<const Ctx*>(h)->stats;
  return VBH_OK;
}

int vbh_spmv(vbh_ctx* h, const double* x, double* y) {
  if (!h) {
    vbh::set_error("vbh_spmv: null ctx");
    return VBH_ERR_ARG;
  }
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded([&] {
    if (!c->have_matrix) vbh::throw_error(VBH_ERR_STATE, "vbh_spmv: no matrix set");
    c->select_stream_device();
    c->ensure_work(c->n_pad);
    c->to_operator(x, false, c->w_r64.p);
    if (c->n_pad > c->n)
      vbh::launch_fill_f64(c->n_pad - c->n, c->w_r64.p + c->n, 0.0, c->stream);
    if (c->format == 0)
      vbh::launch_spmv_csr<double>(c->csr, c->w_r64.p, c->w_Ap.p, c->stream);
    else
      vbh::launch_spmv_bsr<double>(c->bsrd, c->w_r64.p, c->w_Ap.p, c->stream);
    c->from_operator(c->w_Ap.p, y, false);
    VBH_HIP(hipStreamSynchronize(c->stream));
  });
}

int vbh_bench_spmv(vbh_ctx* h, int precision, int reps, double* ms_per_launch) {
  if (!h || reps <= 0 || reps > 1000 || !ms_per_launch || precision < 0 || precision > 14 ||
      precision == 11 || precision == 12) {
    vbh::set_error("vbh_bench_spmv: bad argument");
    return VBH_ERR_ARG;
  }
  Ctx* c = reinterpret_cast<Ctx*>(h);
  return guarded([&] {
    if (!c->have_matrix) vbh::throw_error(VBH_ERR_STATE, "vbh_bench_spmv: no matrix set");
    c->select_stream_device();
    c->ensure_work(c->n_pad);
    vbh::launch_fill_f64(c->n_pad, c->w_r64.p, 1.0, c->stream);
    vbh::launch_cast_f64_to_f32(c->n_pad, c->w_r64.p, c->w_p32.p, c->stream);
    const vbh::CgRed rd = c->red();
    const vbh::PcgParams prm = c->params();
    if (precision >= 2) {  // fused solver kernels need an initialised state (F in w_F)
      if ((precision < 13 && precision % 2 == 0) || precision == 13) {
        vbh::launch_cg_init<double>(c->n_pad, c->bufs64(), c->st_dev.p, rd, c->stream);
        vbh::launch_cg_spmv<double>(c->format, c->csr, c->bsrd, c->bufs64(), c->st_dev.p, rd,
                                    c->stream, 0, 0);
      } else {
        vbh::launch_cg_init<float>(c->n_pad, c->bufs32(), c->st_dev.p, rd, c->stream);
        vbh::launch_cg_spmv<float>(c->format, c->csr, c->bsrd, c->bufs32(), c->st_dev.p, rd,
                                   c->stream, c->spmv_acc64, 0);
      }
    }
    // tol 0, no iteration limit and no refresh pauses: timed solver kernels
    // never find the state stopped
    vbh::PcgParams run_all = prm;
    run_all.tol = 0.0;
    run_all.max_iters = 1 << 30;
    int pos = 0;  // position of the next vector kernel in its launch sequence
    DBuf<char> cp_a, cp_b;
    const long long cp_bytes = 256ll << 20;
    if (precision == 6) {
      cp_a.resize(static_cast<size_t>(cp_bytes));
      cp_b.resize(static_cast<size_t>(cp_bytes));
      VBH_HIP(hipMemsetAsync(cp_a.p, 0, static_cast<size_t>(cp_bytes), c->stream));
    }
    auto one = [&] {
      if (precision == 6) {
        vbh::launch_stream_copy(cp_bytes, cp_a.p, cp_b.p, c->stream);
      } else if (precision == 2) {
        vbh::launch_cg_spmv<double>(c->format, c->csr, c->bsrd, c->bufs64(), c->st_dev.p, rd,
                                    c->stream, 0, 0);
      } else if (precision == 3) {
        vbh::launch_cg_spmv<float>(c->format, c->csr, c->bsrd, c->bufs32(), c->st_dev.p, rd,
                                   c->stream, c->spmv_acc64, 0);
      } else if (precision == 4) {
        vbh::launch_cg_vector<double>(c->n_pad, c->bufs64(), c->st_dev.p, rd, run_all, 0, pos++,
                                      c->stream);
      } else if (precision == 5) {
        vbh::launch_cg_vector<float>(c->n_pad, c->bufs32(), c->st_dev.p, rd, run_all, 0, pos++,
                                     c->stream);
      } else if (precision == 0) {
        if (c->format == 0) vbh::launch_spmv_csr<double>(c->csr, c->w_r64.p, c->w_Ap.p, c->stream);
        else vbh::launch_spmv_bsr<double>(c->bsrd, c->w_r64.p, c->w_Ap.p, c->stream);
      } else {
        if (c->format == 0) vbh::launch_spmv_csr<float>(c->csr, c->w_p32.p, c->w_Ap32.p, c->stream);
        else vbh::launch_spmv_bsr<float>(c->bsrd, c->w_p32.p, c->w_Ap32.p, c->stream);
      }
    };
    if (precision >= 7 && precision <= 10) {
      // In context: `reps` PCG iterations (SpMV, vector kernel) launched
      // directly; the timed kernel goes out through hipExtLaunchKernelGGL
      // with a start/stop event pair, which records that dispatch's own begin
      // and end, so the mean is the kernel duration a profiler reports, with
      // p just written by the vector kernel as in a solve.
      const bool f32 = (precision % 2) == 0;
      const int kind = (precision - 7) / 2;  // 0 SpMV, 1 vector kernel
      std::vector<hipEvent_t> ev(2 * static_cast<size_t>(reps), nullptr);
      for (auto& e : ev) VBH_HIP(hipEventCreate(&e));
      if (f32) vbh::launch_cg_init<float>(c->n_pad, c->bufs32(), c->st_dev.p, rd, c->stream);
      else vbh::launch_cg_init<double>(c->n_pad, c->bufs64(), c->st_dev.p, rd, c->stream);
      for (int i = 0; i < reps + 1; ++i) {  // iteration 0 warms up
        hipEvent_t e0 = i ? ev[2 * static_cast<size_t>(i - 1)] : nullptr;
        hipEvent_t e1 = i ? ev[2 * static_cast<size_t>(i - 1) + 1] : nullptr;
        hipEvent_t s0 = kind == 0 ? e0 : nullptr, s1 = kind == 0 ? e1 : nullptr;
        hipEvent_t v0 = kind == 1 ? e0 : nullptr, v1 = kind == 1 ? e1 : nullptr;
        if (f32) {
          vbh::launch_cg_spmv<float>(c->format, c->csr, c->bsrd, c->bufs32(), c->st_dev.p, rd,
                                     c->stream, c->spmv_acc64, i, s0, s1);
          vbh::launch_cg_vector<float>(c->n_pad, c->bufs32(), c->st_dev.p, rd, run_all, 0, i,
                                       c->stream, v0, v1);
        } else {
          vbh::launch_cg_spmv<double>(c->format, c->csr, c->bsrd, c->bufs64(), c->st_dev.p, rd,
                                      c->stream, 0, i, s0, s1);
          vbh::launch_cg_vector<double>(c->n_pad, c->bufs64(), c->st_dev.p, rd, run_all, 0, i,
                                        c->stream, v0, v1);
        }
      }
      VBH_HIP(hipStreamSynchronize(c->stream));
      double total = 0.0;
      for (int i = 0; i < reps; ++i) {
        float ms = 0.0f;
        VBH_HIP(hipEventElapsedTime(&ms, ev[2 * static_cast<size_t>(i)], ev[2 * static_cast<size_t>(i) + 1]));
        total += ms;
      }
      for (auto& e : ev) (void)hipEventDestroy(e);
      vbh::PcgScalars st{};
      VBH_HIP(hipMemcpyAsync(&st, c->st_dev.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
      VBH_HIP(hipStreamSynchronize(c->stream));
      if (st.done)
        vbh::throw_error(VBH_ERR_SOLVER, "vbh_bench_spmv: the PCG state stopped during timing");
      *ms_per_launch = total / reps;
      return;
    }
    if (precision == 13 || precision == 14) {
      // the fused SpMV launched `reps` times back to back between ONE event
      // pair: the mean holds the kernel plus its share of the dispatch gaps,
      // and none of the per-launch event overhead
      const int w = precision == 13 ? 2 : 3;
      for (int i = 0; i < 2; ++i) {  // warm up
        if (w == 2) vbh::launch_cg_spmv<double>(c->format, c->csr, c->bsrd, c->bufs64(), c->st_dev.p, rd, c->stream, 0, 0);
        else vbh::launch_cg_spmv<float>(c->format, c->csr, c->bsrd, c->bufs32(), c->st_dev.p, rd, c->stream, c->spmv_acc64, 0);
      }
      hipEvent_t e0 = nullptr, e1 = nullptr;
      VBH_HIP(hipEventCreate(&e0));
      VBH_HIP(hipEventCreate(&e1));
      VBH_HIP(hipEventRecord(e0, c->stream));
      for (int i = 0; i < reps; ++i) {
        if (w == 2) vbh::launch_cg_spmv<double>(c->format, c->csr, c->bsrd, c->bufs64(), c->st_dev.p, rd, c->stream, 0, 0);
        else vbh::launch_cg_spmv<float>(c->format, c->csr, c->bsrd, c->bufs32(), c->st_dev.p, rd, c->stream, c->spmv_acc64, 0);
      }
      VBH_HIP(hipEventRecord(e1, c->stream));
      VBH_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      VBH_HIP(hipEventElapsedTime(&ms, e0, e1));
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      *ms_per_launch = static_cast<double>(ms) / reps;
      return;
    }
    one();
    // one event pair around every launch (includes the event overhead)
    std::vector<hipEvent_t> ev(2 * static_cast<size_t>(reps), nullptr);
    for (auto& e : ev) VBH_HIP(hipEventCreate(&e));
    for (int i = 0; i < reps; ++i) {
      VBH_HIP(hipEventRecord(ev[2 * static_cast<size_t>(i)], c->stream));
      one();
      VBH_HIP(hipEventRecord(ev[2 * static_cast<size_t>(i) + 1], c->stream));
    }
    VBH_HIP(hipEventSynchronize(ev.back()));
    double total = 0.0;
    for (int i = 0; i < reps; ++i) {
      float ms = 0.0f;
      VBH_HIP(hipEventElapsedTime(&ms, ev[2 * static_cast<size_t>(i)], ev[2 * static_cast<size_t>(i) + 1]));
      total += ms;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    *ms_per_launch = total / reps;
  });
}

int vbh_operator_stream_bytes(const vbh_ctx* h, int precision, long long* bytes) {
  if (!h || !bytes || (precision != VBH_PREC_FP64 && precision != VBH_PREC_MIXED)) {
    vbh::set_error("vbh_operator_stream_bytes: bad argument");
    return VBH_ERR_ARG;
  }
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  if (!c->have_matrix) {
    vbh::set_error("no matrix set");
    return VBH_ERR_STATE;
  }
  const long long sv = precision == VBH_PREC_FP64 ? 8 : 4;
  if (c->format == 0) {
    *bytes = static_cast<long long>(c->csr.nnz) * (sv + 4) + 4ll * (c->csr.n + 1);
    return VBH_OK;
  }
  const vbh::BsrDev& A = c->bsrd;
  const long long B = A.B;
  if (A.rk64) {  // relative sliced ELL: values, one 32-bit offset per slot, slice offsets
    const long long slots = precision == VBH_PREC_FP64 ? c->rel_slots64 : c->rel_slots32;
    long long b = slots * vbh::kSliceRows * B * sv + slots * 4 + 4ll * (A.nslices + 1);
    if (precision == VBH_PREC_MIXED && A.dblk) b += static_cast<long long>(A.nbr) * B * B * 8;
    *bytes = b;
    return VBH_OK;
  }
  const long long slots = c->sell_slots;
  long long b = slots * vbh::kSliceRows * B * sv;                        // values
  b += slots * vbh::sell_nps(A.B) * (A.sidx16 ? 2 : 4);                 // block columns
  b += 4ll * (A.nslices + 1);                                           // slice offsets
  if (precision == VBH_PREC_MIXED && A.dblk) b += static_cast<long long>(A.nbr) * B * B * 8;
  *bytes = b;
  return VBH_OK;
}

int vbh_device_alloc(vbh_ctx* h, long long bytes, void** dptr) {
  if (!h || !dptr || bytes < 0) {
    vbh::set_error("vbh_device_alloc: bad argument");
    return VBH_ERR_ARG;
  }
  return guarded([&] {
    reinterpret_cast<Ctx*>(h)->select_stream_device();
    VBH_HIP(hipMalloc(dptr, static_cast<size_t>(bytes > 0 ? bytes : 1)));
  });
}

int vbh_device_free(vbh_ctx* h, void* dptr) {
  (void)h;
  return guarded([&] { VBH_HIP(hipFree(dptr)); });
}

int vbh_memcpy_h2d(vbh_ctx* h, void* dst, const void* src, long long bytes) {
  return guarded([&] {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    VBH_HIP(hipMemcpyAsync(dst, src, static_cast<size_t>(bytes), hipMemcpyHostToDevice, c->stream));
    VBH_HIP(hipStreamSynchronize(c->stream));
  });
}

int vbh_memcpy_d2h(vbh_ctx* h, void* dst, const void* src, long long bytes) {
  return guarded([&] {
    Ctx* c = reinterpret_cast<Ctx*>(h);
    VBH_HIP(hipMemcpyAsync(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToHost, c->stream));
    VBH_HIP(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"

extern "C" {
// Block-PCG entry points: implemented in vbh_block.hip.
}
