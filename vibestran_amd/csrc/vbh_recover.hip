// vibestran_amd/csrc/vbh_recover.hip -- result recovery on the device and the
// CSV result files (gfx950 kernels, host writer).
//
// Replaces the element loop of LinearStaticSolver::recover_results
// (reference src/solver/linear_static.cpp:566-1201) over the mesh kept by the
// last vbh_assemble:
//   displacements  per node, T1..R3, 0 for constrained or inactive DOFs
//                  (568-612; basic frame, no CD rotation)
//   solids         sigma = D (B u) with the compatible B of the point
//                  (solid_sigma_from_B 186-196). EAS elements are recovered
//                  with that displacement B too: their enhanced modes are
//                  condensed out of the stiffness and recovery never sees
//                  them. CHEXA8: centroid and the 8 corners (1058-1117,
//                  1167-1190); CPENTA6: (1/3, 1/3, 0) and the 6 corners
//                  (996-1056, 1146-1166); CTETRA4: the constant CST stress at
//                  the centroid and every node (830-882, 1120-1124)
//   CQUAD4         membrane stress and moments at the centroid and the four
//                  corners (712-753; kernel in vbh_shell.hip)
// One thread per element: recovery is a handful of small products per point
// against ~1 KB of gathered inputs and outputs per element, so the kernels are
// bound by the gathers and the row stores, not by arithmetic.
//
// The CSV writer reproduces CsvWriter::write (src/io/csv_writer.cpp:21-93):
// "%.8e" values (std::scientific, setprecision(8)), ", " separators, one
// header line, node rows in ascending node ID order, element rows plates first
// and solids after.
#include <algorithm>
#include <cstdio>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "vbh_ctx.hpp"
#include "vbh_elem.hpp"

namespace vbh {
namespace {

// Parametric shape derivatives d[n][k] = dN_n/dr_k: CHexa8::shape_functions
// (solid_elements.cpp:76-93), CPenta6::shape_functions (510-548) with
// (L1, L2, zeta), and the CTETRA4 barycentric coordinates (L1, L2, L3) of
// nodes 1..3 (node 0 takes 1 - L1 - L2 - L3).
template <int NN>
__device__ __forceinline__ void param_d(double r0, double r1, double r2, double (&d)[NN][3]) {
  if constexpr (NN == 8) {
    constexpr double xn[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    constexpr double en[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    constexpr double zn[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      d[i][0] = 0.125 * xn[i] * (1 + en[i] * r1) * (1 + zn[i] * r2);
      d[i][1] = 0.125 * en[i] * (1 + xn[i] * r0) * (1 + zn[i] * r2);
      d[i][2] = 0.125 * zn[i] * (1 + xn[i] * r0) * (1 + en[i] * r1);
    }
  } else if constexpr (NN == 6) {
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
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) d[i][k] = (i == 0) ? -1.0 : (i == k + 1 ? 1.0 : 0.0);
  }
}

// Stress row at one parametric point:
// [sx, sy, sxy, sz, syz, szx, 0, 0, 0, von_mises] (csv_writer.cpp:76-80 order;
// solid_von_mises linear_static.cpp:140-151).
template <int NN>
__device__ void solid_point_row(const double (&X)[NN][3], const double (&ue)[3 * NN], double lam,
                                double mu, double r0, double r1, double r2, double* out,
                                int& bad) {
  double d[NN][3];
  param_d<NN>(r0, r1, r2, d);
  double J[3][3] = {};
#pragma unroll
  for (int n = 0; n < NN; ++n)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) J[i][j] += d[n][i] * X[n][j];
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c10 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c20 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double det = J[0][0] * c00 + J[0][1] * c10 + J[0][2] * c20;
  if (det == 0.0) bad |= kErrJacobian;
  const double id = 1.0 / det;
  const double Ji[3][3] = {
      {c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id,
       (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id},
      {c10 * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id,
       (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id},
      {c20 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id,
       (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
  // displacement gradient H[a][i] = d u_a / d x_i
  double H[3][3] = {};
#pragma unroll
  for (int n = 0; n < NN; ++n) {
    double g[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) g[i] = Ji[i][0] * d[n][0] + Ji[i][1] * d[n][1] + Ji[i][2] * d[n][2];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) H[a][i] += g[i] * ue[3 * n + a];
  }
  const double tr = H[0][0] + H[1][1] + H[2][2];
  const double s[6] = {lam * tr + 2.0 * mu * H[0][0], lam * tr + 2.0 * mu * H[1][1],
                       lam * tr + 2.0 * mu * H[2][2], mu * (H[0][1] + H[1][0]),
                       mu * (H[1][2] + H[2][1]),      mu * (H[0][2] + H[2][0])};
  const double vm = sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[1] - s[2]) * (s[1] - s[2]) +
                                (s[2] - s[0]) * (s[2] - s[0]) +
                                6.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])));
  const double row[10] = {s[0], s[1], s[3], s[2], s[4], s[5], 0.0, 0.0, 0.0, vm};
#pragma unroll
  for (int k = 0; k < 10; ++k) out[k] = row[k];
}

template <int NN>
__global__ __launch_bounds__(256) void k_recover_solid(
    int nsel, const int* __restrict__ sel, const int* __restrict__ eptr,
    const int* __restrict__ conn, const double* __restrict__ xyz, const double* __restrict__ E,
    const double* __restrict__ nu, const int* __restrict__ gd, const double* __restrict__ u,
    double* __restrict__ cen, double* __restrict__ cor, int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nsel) return;
  const int e = sel[i];
  const int o = eptr[e];
  double X[NN][3], ue[3 * NN];
  const int* g = gd + static_cast<size_t>(e) * kGdStride;
#pragma unroll
  for (int a = 0; a < NN; ++a) {
    const int node = conn[o + a];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      X[a][d] = xyz[3 * node + d];
      const int q = g[3 * a + d];
      ue[3 * a + d] = q >= 0 ? u[q] : 0.0;
    }
  }
  // isotropic D (solid_constitutive_matrix, linear_static.cpp:119-138)
  const double Ee = E[e], ne = nu[e];
  const double lam = Ee * ne / ((1.0 + ne) * (1.0 - 2.0 * ne));
  const double mu = Ee / (2.0 * (1.0 + ne));
  int bad = 0;
  double* c = cen + 10 * static_cast<size_t>(e);
  if constexpr (NN == 8) {
    solid_point_row<8>(X, ue, lam, mu, 0.0, 0.0, 0.0, c, bad);
  } else if constexpr (NN == 6) {
    solid_point_row<6>(X, ue, lam, mu, 1.0 / 3.0, 1.0 / 3.0, 0.0, c, bad);
  } else {
    solid_point_row<4>(X, ue, lam, mu, 0.25, 0.25, 0.25, c, bad);
  }
  if (cor) {
    double* r = cor + 80 * static_cast<size_t>(e);
    if constexpr (NN == 8) {
      constexpr double xn[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
      constexpr double en[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
      constexpr double zn[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
      for (int k = 0; k < 8; ++k)
        solid_point_row<8>(X, ue, lam, mu, xn[k], en[k], zn[k], r + 10 * k, bad);
    } else if constexpr (NN == 6) {
      constexpr double l1[6] = {1, 0, 0, 1, 0, 0};
      constexpr double l2[6] = {0, 1, 0, 0, 1, 0};
      constexpr double zz[6] = {-1, -1, -1, 1, 1, 1};
      for (int k = 0; k < 6; ++k)
        solid_point_row<6>(X, ue, lam, mu, l1[k], l2[k], zz[k], r + 10 * k, bad);
      for (int k = 60; k < 80; ++k) r[k] = 0.0;
    } else {
      for (int k = 0; k < 4; ++k)
        for (int j = 0; j < 10; ++j) r[10 * k + j] = c[j];
      for (int k = 40; k < 80; ++k) r[k] = 0.0;
    }
  }
  if (bad) {
    atomicOr(&err[0], bad);
    atomicMin(&err[1], e);
  }
}

__global__ void k_node_displacements(int m, const int* __restrict__ eq,
                                     const double* __restrict__ u, double* __restrict__ disp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int q = eq[i];
  disp[i] = q >= 0 ? u[q] : 0.0;
}

__global__ void k_recover_err_init(int* err) {
  err[0] = 0;
  err[1] = 0x7fffffff;
}

}  // namespace

void Ctx::recover_stress(const double* u, double* centroid, double* corner, bool dev) {
  if (!mesh.valid) throw_error(VBH_ERR_STATE, "vbh_recover_stress: no assembled mesh in this context");
  if (!u || !centroid) throw_error(VBH_ERR_ARG, "vbh_recover_stress: null u or centroid");
  select_stream_device();
  hipStream_t s = stream;
  const size_t ne = static_cast<size_t>(mesh.nelem);
  DBuf<double> du, fu, dcen, dcor;
  DBuf<int> derr;
  const double* pu = mesh_solution(u, dev, du, fu);
  double* pc = centroid;
  double* pk = corner;
  if (!dev) {
    dcen.resize(10 * ne);
    pc = dcen.p;
    if (corner) {
      dcor.resize(80 * ne);
      pk = dcor.p;
    }
  }
  derr.resize(2);
  hipLaunchKernelGGL(k_recover_err_init, dim3(1), dim3(1), 0, s, derr.p);
  // element groups of the retained mesh: 0 CHEXA8 EAS, 1 CHEXA8 SRI,
  // 2 CPENTA6, 3 CTETRA4, 4 CQUAD4
  for (int k = 0; k < 5; ++k) {
    const int ns = mesh.nsel[k];
    if (!ns) continue;
    const dim3 g((ns + 255) / 256), b(256);
    if (k == 4)
      launch_quad4_recover(ns, mesh.sel[k].p, mesh.ep.p, mesh.conn.p, mesh.xyz.p, mesh.E.p,
                           mesh.nu.p, mesh.t.p, mesh.shell_form, mesh.gd.p, pu, pc, pk, derr.p, s);
    else if (k <= 1)
      hipLaunchKernelGGL(k_recover_solid<8>, g, b, 0, s, ns, mesh.sel[k].p, mesh.ep.p,
                         mesh.conn.p, mesh.xyz.p, mesh.E.p, mesh.nu.p, mesh.gd.p, pu, pc, pk,
                         derr.p);
    else if (k == 2)
      hipLaunchKernelGGL(k_recover_solid<6>, g, b, 0, s, ns, mesh.sel[k].p, mesh.ep.p,
                         mesh.conn.p, mesh.xyz.p, mesh.E.p, mesh.nu.p, mesh.gd.p, pu, pc, pk,
                         derr.p);
    else
      hipLaunchKernelGGL(k_recover_solid<4>, g, b, 0, s, ns, mesh.sel[k].p, mesh.ep.p,
                         mesh.conn.p, mesh.xyz.p, mesh.E.p, mesh.nu.p, mesh.gd.p, pu, pc, pk,
                         derr.p);
  }
  VBH_HIP(hipGetLastError());
  int herr[2];
  VBH_HIP(hipMemcpyAsync(herr, derr.p, sizeof herr, hipMemcpyDeviceToHost, s));
  if (!dev) {
    VBH_HIP(hipMemcpyAsync(centroid, dcen.p, sizeof(double) * 10 * ne, hipMemcpyDeviceToHost, s));
    if (corner)
      VBH_HIP(hipMemcpyAsync(corner, dcor.p, sizeof(double) * 80 * ne, hipMemcpyDeviceToHost, s));
  }
  VBH_HIP(hipStreamSynchronize(s));
  if (herr[0])
    throw_error(VBH_ERR_SOLVER, "vbh_recover_stress: element " + std::to_string(herr[1]) +
                                    ": singular Jacobian or MITC4+ geometry (code " +
                                    std::to_string(herr[0]) + ")");
}

void Ctx::recover_displacements(const double* u, double* disp, bool dev) {
  if (!mesh.valid)
    throw_error(VBH_ERR_STATE, "vbh_recover_displacements: no assembled mesh in this context");
  if (!u || !disp) throw_error(VBH_ERR_ARG, "vbh_recover_displacements: null u or disp");
  select_stream_device();
  hipStream_t s = stream;
  const int m = 6 * mesh.nnodes;
  DBuf<double> du, fu, dd;
  const double* pu = mesh_solution(u, dev, du, fu);
  double* pd = disp;
  if (!dev) {
    dd.resize(static_cast<size_t>(m));
    pd = dd.p;
  }
  if (m > 0)
    hipLaunchKernelGGL(k_node_displacements, dim3((m + 255) / 256), dim3(256), 0, s, m,
                       mesh.eq.p, pu, pd);
  VBH_HIP(hipGetLastError());
  if (!dev) VBH_HIP(hipMemcpyAsync(disp, dd.p, sizeof(double) * m, hipMemcpyDeviceToHost, s));
  VBH_HIP(hipStreamSynchronize(s));
}

namespace {

// Rows [lo, hi) formatted by `fmt_row` into one buffer per worker, then written
// in order.
template <class Fmt>
void write_rows(FILE* f, long long nrows, Fmt fmt_row) {
  const long long per = 65536;
  const long long nchunk = (nrows + per - 1) / per;
  const int nthreads = static_cast<int>(std::min<long long>(
      std::max(1u, std::min(16u, std::thread::hardware_concurrency())), std::max(1LL, nchunk)));
  for (long long base = 0; base < nchunk; base += nthreads) {
    const int nt = static_cast<int>(std::min<long long>(nthreads, nchunk - base));
    std::vector<std::string> out(static_cast<size_t>(nt));
    std::vector<std::thread> pool;
    auto work = [&](int t) {
      const long long lo = (base + t) * per, hi = std::min(nrows, lo + per);
      std::string& b = out[static_cast<size_t>(t)];
      b.reserve(static_cast<size_t>(hi - lo) * 200);
      char line[512];
      for (long long r = lo; r < hi; ++r) {
        const int len = fmt_row(r, line, sizeof line);
        b.append(line, static_cast<size_t>(len));
      }
    };
    for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    for (const auto& b : out)
      if (std::fwrite(b.data(), 1, b.size(), f) != b.size())
        throw_error(VBH_ERR_SOLVER, "CSV write failed");
  }
}

FILE* open_csv(const char* path, int append, const char* header) {
  if (!path) throw_error(VBH_ERR_ARG, "CSV: null path");
  FILE* f = std::fopen(path, append ? "ab" : "wb");
  if (!f) throw_error(VBH_ERR_SOLVER, std::string("Cannot write CSV: ") + path);
  if (!append) std::fputs(header, f);
  return f;
}

void close_csv(FILE* f, const char* path) {
  if (std::fclose(f) != 0) throw_error(VBH_ERR_SOLVER, std::string("Cannot write CSV: ") + path);
}

}  // namespace

}  // namespace vbh

extern "C" {

int vbh_recover_stress(vbh_ctx* h, const double* u, double* centroid, double* corner) {
  if (!h) {
    vbh::set_error("vbh_recover_stress: null ctx");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] { c->recover_stress(u, centroid, corner, false); });
}

int vbh_recover_stress_device(vbh_ctx* h, const double* du, double* dcentroid, double* dcorner) {
  if (!h) {
    vbh::set_error("vbh_recover_stress_device: null ctx");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] { c->recover_stress(du, dcentroid, dcorner, true); });
}

int vbh_recover_displacements(vbh_ctx* h, const double* u, double* disp) {
  if (!h) {
    vbh::set_error("vbh_recover_displacements: null ctx");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] { c->recover_displacements(u, disp, false); });
}

int vbh_recover_displacements_device(vbh_ctx* h, const double* du, double* ddisp) {
  if (!h) {
    vbh::set_error("vbh_recover_displacements_device: null ctx");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] { c->recover_displacements(du, ddisp, true); });
}

int vbh_write_node_csv(const char* path, int append, int subcase_id, int nnodes,
                       const int* node_ids, const double* disp) {
  return vbh::guarded([&] {
    if (nnodes < 0 || (nnodes > 0 && (!node_ids || !disp)))
      vbh::throw_error(VBH_ERR_ARG, "vbh_write_node_csv: bad node arrays");
    std::vector<int> order(static_cast<size_t>(nnodes));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return node_ids[a] < node_ids[b]; });
    FILE* f = vbh::open_csv(path, append, "# node_id, subcase_id, T1, T2, T3, R1, R2, R3\n");
    try {
      vbh::write_rows(f, nnodes, [&](long long r, char* line, size_t cap) {
        const int i = order[static_cast<size_t>(r)];
        const double* d = disp + 6 * static_cast<size_t>(i);
        return std::snprintf(line, cap, "%d, %d, %.8e, %.8e, %.8e, %.8e, %.8e, %.8e\n",
                             node_ids[i], subcase_id, d[0], d[1], d[2], d[3], d[4], d[5]);
      });
    } catch (...) {
      std::fclose(f);
      throw;
    }
    vbh::close_csv(f, path);
  });
}

int vbh_write_elem_csv(const char* path, int append, int subcase_id, int nelem,
                       const int* elem_ids, const int* etype, const double* rows) {
  return vbh::guarded([&] {
    if (nelem < 0 || (nelem > 0 && (!elem_ids || !etype || !rows)))
      vbh::throw_error(VBH_ERR_ARG, "vbh_write_elem_csv: bad element arrays");
    // plate rows first, then solid rows, each in element order
    std::vector<int> order;
    order.reserve(static_cast<size_t>(nelem));
    for (int e = 0; e < nelem; ++e)
      if (etype[e] == VBH_CQUAD4) order.push_back(e);
    for (int e = 0; e < nelem; ++e)
      if (etype[e] == VBH_CHEXA8 || etype[e] == VBH_CPENTA6 || etype[e] == VBH_CTETRA4)
        order.push_back(e);
    FILE* f = vbh::open_csv(path, append,
                            "# elem_id, subcase_id, sx, sy, sxy, sz, syz, szx,"
                            " mx, my, mxy, von_mises\n");
    try {
      vbh::write_rows(f, static_cast<long long>(order.size()),
                      [&](long long r, char* line, size_t cap) {
                        const int e = order[static_cast<size_t>(r)];
                        const double* v = rows + 10 * static_cast<size_t>(e);
                        return std::snprintf(
                            line, cap,
                            "%d, %d, %.8e, %.8e, %.8e, %.8e, %.8e, %.8e, %.8e, %.8e, %.8e, %.8e\n",
                            elem_ids[e], subcase_id, v[0], v[1], v[2], v[3], v[4], v[5], v[6],
                            v[7], v[8], v[9]);
                      });
    } catch (...) {
      std::fclose(f);
      throw;
    }
    vbh::close_csv(f, path);
  });
}

}  // extern "C"
