// vibestran_amd/csrc/vbh_noalloc.hip -- the single-load-case Jacobi-PCG solve
// without any device allocation (vbh_solve_device_noalloc).
//
// For callers that cannot allocate device memory during a solve (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
// every work buffer of the solve -- the PCG state scalars, the per-block
// partial sums and the F, u, r, z, p, Ap (and FP64 refresh residual) vectors
// -- lives in one file-scope device array, `scratch`, of kScratchBytes per
// GPU. The operator itself is set up beforehand (vbh_set_csr / vbh_assemble),
// which may allocate; this entry point then calls no hipMalloc, hipFree or
// other allocating routine, and launches the solver kernels directly (no
// graph capture or instantiation).
//
// The array's name is a host-side stand-in, not a device address: at the
// library's set-up call (vbh_ctx_create, with the context's device current)
// scratch_init() takes its device address with hipGetSymbolAddress, once per
// device, and keeps it; only that address reaches kernels and copies.
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (two host threads, or two contexts' streams) race on it.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <mutex>
#include <vector>

#include "vbh_ctx.hpp"

namespace vbh {

namespace {

constexpr size_t kScratchBytes = size_t(96) << 20;  // 96 MiB per GPU
constexpr size_t kAlign = 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_ptr;  // device address of `scratch`, per device

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Carves the scratch array into the buffers of one solve of n_pad rows.
struct ScratchLayout {
  size_t st, partA, partB, F, u, r64, r, p, Ap, end;
  ScratchLayout(size_t n_pad, bool mixed) {
    const size_t part = sizeof(double) * kMaxRed * kMaxGrid;
    const size_t vs = mixed ? sizeof(float) : sizeof(double);
    size_t o = 0;
    st = o;
    o = align_up(o + sizeof(PcgScalars));
    partA = o;
    o = align_up(o + part);
    partB = o;
    o = align_up(o + 2 * part);
    F = o;
    o = align_up(o + sizeof(double) * n_pad);
    u = o;
    o = align_up(o + sizeof(double) * n_pad);
    r64 = o;
    if (mixed) o = align_up(o + sizeof(double) * n_pad);
    r = o;
    o = align_up(o + vs * n_pad);
    p = o;
    o = align_up(o + vs * n_pad);
    Ap = o;
    o = align_up(o + vs * n_pad);
    end = o;
  }
};

}  // namespace

// Takes the device address of `scratch` on `device` (current) once per device.
void scratch_init(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0) throw_error(VBH_ERR_ARG, "scratch_init: bad device");
  if (g_scratch_ptr.size() <= static_cast<size_t>(device))
    g_scratch_ptr.resize(static_cast<size_t>(device) + 1, nullptr);
  if (g_scratch_ptr[static_cast<size_t>(device)]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    throw_error(VBH_ERR_HIP, std::string("HIP error '") + hipGetErrorString(e) +
                                 "' at hipGetSymbolAddress(scratch)");
  g_scratch_ptr[static_cast<size_t>(device)] = static_cast<unsigned char*>(ptr);
}

static unsigned char* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || static_cast<size_t>(device) >= g_scratch_ptr.size() ||
      !g_scratch_ptr[static_cast<size_t>(device)])
    throw_error(VBH_ERR_STATE, "vbh_solve_device_noalloc: scratch of device " +
                                   std::to_string(device) + " not set up");
  return g_scratch_ptr[static_cast<size_t>(device)];
}

// The solve of Ctx::run_solve on scratch buffers, iterations launched directly.
static PcgScalars run_solve_noalloc(Ctx& c, int precision_mode, const double* dF, double* du) {
  const bool mixed = precision_mode == VBH_PREC_MIXED;
  const size_t N = static_cast<size_t>(c.n_pad);
  const ScratchLayout L(N, mixed);
  if (L.end > kScratchBytes)
    throw_error(VBH_ERR_ARG,
                "vbh_solve_device_noalloc: " + std::to_string(N) + " operator rows need " +
                    std::to_string(L.end) + " bytes of scratch, more than its " +
                    std::to_string(kScratchBytes) + " (at most " +
                    std::to_string((kScratchBytes - L.F) / (mixed ? 36 : 40)) +
                    " rows in this precision)");
  unsigned char* base = scratch_of(c.device);
  PcgScalars* st = reinterpret_cast<PcgScalars*>(base + L.st);
  CgRed rd;
  rd.partA = reinterpret_cast<double*>(base + L.partA);
  rd.partB = reinterpret_cast<double*>(base + L.partB);
  const int rows = (c.format == 0) ? c.n : c.bsrd.nslices;
  rd.nA = grid_for(rows, rows_per_block(c.format));
  rd.nB = vector_grid(static_cast<long long>(N), c.n_cu);
  double* F = reinterpret_cast<double*>(base + L.F);
  double* u = reinterpret_cast<double*>(base + L.u);
  double* r64 = reinterpret_cast<double*>(base + L.r64);
  const CgBufs<double> b64{F, c.diag64.p, reinterpret_cast<double*>(base + L.r),
                           reinterpret_cast<double*>(base + L.p),
                           reinterpret_cast<double*>(base + L.Ap), u, nullptr,
                           c.format != 0 ? c.bsrd.xcd[0] : 0};
  const CgBufs<float> b32{F, c.diag32.p, reinterpret_cast<float*>(base + L.r),
                          reinterpret_cast<float*>(base + L.p), reinterpret_cast<float*>(base + L.Ap),
                          u, r64, c.format != 0 ? c.bsrd.xcd[1] : 0};
  hipStream_t s = c.stream;
  const PcgParams prm = c.params();
  // F in the operator's layout; the padding rows past n stay zero (a fill
  // kernel: the scratch keeps the previous call's contents)
  if (N > static_cast<size_t>(c.n))
    launch_fill_f64(static_cast<long long>(N - c.n), F + c.n, 0.0, s);
  c.to_operator(dF, true, F);
  VBH_HIP(hipEventRecord(c.ev_start, s));
  if (mixed) launch_cg_init<float>(static_cast<int>(N), b32, st, rd, s);
  else launch_cg_init<double>(static_cast<int>(N), b64, st, rd, s);
  const int per = c.iters_per_graph;  // even (see launch_cg_vector)
  const PcgScalars out = c.drive(st, [&] {
    for (int i = 0; i < per; ++i) {
      if (mixed) {
        launch_cg_spmv<float>(c.format, c.csr, c.bsrd, b32, st, rd, s, c.spmv_acc64, i);
        launch_cg_vector<float>(static_cast<int>(N), b32, st, rd, prm, 1, i, s);
      } else {
        launch_cg_spmv<double>(c.format, c.csr, c.bsrd, b64, st, rd, s, 0, i);
        launch_cg_vector<double>(static_cast<int>(N), b64, st, rd, prm, 0, i, s);
      }
    }
  });
  VBH_HIP(hipEventRecord(c.ev_stop, s));
  VBH_HIP(hipEventSynchronize(c.ev_stop));
  float ms = 0.0f;
  VBH_HIP(hipEventElapsedTime(&ms, c.ev_start, c.ev_stop));
  c.from_operator(u, du, true);
  VBH_HIP(hipStreamSynchronize(s));
  c.stats.iterations = out.iter_done;
  c.stats.refinements = out.refreshes;
  c.stats.rel_res = (mixed && (out.flags & (kFlagConverged | kFlagStall))) ? out.rel_true : out.rel;
  if (out.flags & kFlagZeroRhs) c.stats.rel_res = 0.0;
  c.stats.solve_ms = ms;
  c.stats.format = c.format;
  return out;
}

}  // namespace vbh

extern "C" int vbh_solve_device_noalloc(vbh_ctx* h, const double* dF, double* du, int* iters,
                                        double* rel_res) {
  if (!h) {
    vbh::set_error("vbh_solve_device_noalloc: null ctx");
    return VBH_ERR_ARG;
  }
  vbh::Ctx* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] {
    if (!c->have_matrix) vbh::throw_error(VBH_ERR_STATE, "vbh_solve_device_noalloc: no matrix set");
    if (!dF || !du) vbh::throw_error(VBH_ERR_ARG, "vbh_solve_device_noalloc: null vector");
    if (!c->st_host) vbh::throw_error(VBH_ERR_STATE, "vbh_solve_device_noalloc: context not set up");
    c->select_stream_device();
    const vbh::PcgScalars s = vbh::run_solve_noalloc(*c, c->precision, dF, du);
    if (iters) *iters = s.iter_done;
    if (rel_res) *rel_res = c->stats.rel_res;
    c->check_outcome(s, c->precision);
  });
}
