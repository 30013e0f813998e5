// vibestran_amd/csrc/vbh_rhs.hip -- load vector of one load case on the device,
// over the mesh kept by the last vbh_assemble (gfx950).
//
// Replaces the reference's host load loops of LinearStaticSolver:
//   FORCE / MOMENT      linear_static.cpp:344-398   (nodal components, basic frame)
//   PLOAD4 on solids    linear_static_pressure.cpp:154-196 (2x2 Gauss on the
//                       face, normal_sign -1 for solids, 745-753); the face
//                       choice and its outward corner order (420-517) are made
//                       by the caller
//   GRAV                linear_static.cpp:400-503: f_e = M_e a, M_e the
//                       consistent mass. With one acceleration for every
//                       node the translational rows of M_e a reduce to
//                       rho a int(N_a) (the N_b sum to one), evaluated with the
//                       mass matrix's own quadrature: CHEXA8 2x2x2,
//                       CPENTA6 3x2 points of weight 1/6, CTETRA4 rho V/4,
//                       CQUAD4 rho t int(N_a) dA (solid_elements.cpp:182-217,
//                       386-411, 1469-1509; cquad4.cpp:672-708). Elements whose
//                       force is below 1e-30 in every entry add nothing, as in
//                       the reference (495-497).
// Contributions are added with atomics (summation order varies, as the
// reference's element order differs from ours only by rounding).
#include <algorithm>
#include <cmath>
#include <vector>

#include "vbh_ctx.hpp"
#include "vbh_elem.hpp"

namespace vbh {
namespace {

constexpr double kG3 = 0.57735026918962584;  // 1/sqrt(3)

__global__ void k_rhs_nodal(int nnodes, const int* __restrict__ eq,
                            const double* __restrict__ nodal, double* __restrict__ F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 6 * nnodes) return;
  const int q = eq[i];
  const double v = nodal[i];
  if (q >= 0 && v != 0.0) atomicAdd(&F[q], v);
}

// One face per thread: corners c[4] (outward order), corner pressures p[4].
__global__ void k_rhs_pload4(int nface, const int* __restrict__ fn, const double* __restrict__ fp,
                             const double* __restrict__ xyz, const int* __restrict__ eq,
                             double* __restrict__ F) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nface) return;
  double X[4][3], p[4];
  int c[4];
  for (int a = 0; a < 4; ++a) {
    c[a] = fn[4 * f + a];
    p[a] = fp[4 * f + a];
    for (int d = 0; d < 3; ++d) X[a][d] = xyz[3 * c[a] + d];
  }
  double out[4][3] = {};
  for (int gi = 0; gi < 2; ++gi)
    for (int gj = 0; gj < 2; ++gj) {
      const double xi = gi ? kG3 : -kG3, eta = gj ? kG3 : -kG3;
      const double N[4] = {0.25 * (1 - xi) * (1 - eta), 0.25 * (1 + xi) * (1 - eta),
                           0.25 * (1 + xi) * (1 + eta), 0.25 * (1 - xi) * (1 + eta)};
      const double dxi[4] = {-0.25 * (1 - eta), 0.25 * (1 - eta), 0.25 * (1 + eta), -0.25 * (1 + eta)};
      const double det[4] = {-0.25 * (1 - xi), -0.25 * (1 + xi), 0.25 * (1 + xi), 0.25 * (1 - xi)};
      double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
      double pq = 0.0;
      for (int k = 0; k < 4; ++k) {
        pq += N[k] * p[k];
        for (int d = 0; d < 3; ++d) {
          a[d] += dxi[k] * X[k][d];
          b[d] += det[k] * X[k][d];
        }
      }
      const double av[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                            a[0] * b[1] - a[1] * b[0]};
      for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) out[k][d] += -pq * av[d] * N[k];  // solids: -p n dA
    }
  for (int k = 0; k < 4; ++k)
    for (int d = 0; d < 3; ++d) {
      const int q = eq[6 * c[k] + d];
      if (q >= 0) atomicAdd(&F[q], out[k][d]);
    }
}

__device__ __forceinline__ double det3v(const double (&J)[3][3]) {
  return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
         J[1][0] * (J[0][1] * J[2][2] - J[0][2] * J[2][1]) +
         J[2][0] * (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
}

// rho * int(N_a) of one element with its mass quadrature; ok=false on a
// non-positive Jacobian.
__device__ void solid_shape_integrals(int et, const double (&X)[8][3], double rho, double (&w)[8],
                                      bool& ok) {
  for (int a = 0; a < 8; ++a) w[a] = 0.0;
  if (et == VBH_CTETRA4) {
    double m[3][3];
    for (int r = 0; r < 3; ++r)
      for (int d = 0; d < 3; ++d) m[r][d] = X[r + 1][d] - X[0][d];
    const double V = fabs(det3v(m)) / 6.0;
    for (int a = 0; a < 4; ++a) w[a] = rho * V / 4.0;
    return;
  }
  if (et == VBH_CHEXA8) {
    const double xn[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    const double en[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    const double zn[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    for (int g = 0; g < 8; ++g) {
      const double r0 = (g & 4) ? kG3 : -kG3, r1 = (g & 2) ? kG3 : -kG3, r2 = (g & 1) ? kG3 : -kG3;
      double J[3][3] = {};
      double N[8];
      for (int i = 0; i < 8; ++i) {
        N[i] = 0.125 * (1 + xn[i] * r0) * (1 + en[i] * r1) * (1 + zn[i] * r2);
        const double d0 = 0.125 * xn[i] * (1 + en[i] * r1) * (1 + zn[i] * r2);
        const double d1 = 0.125 * en[i] * (1 + xn[i] * r0) * (1 + zn[i] * r2);
        const double d2 = 0.125 * zn[i] * (1 + xn[i] * r0) * (1 + en[i] * r1);
        for (int d = 0; d < 3; ++d) {
          J[0][d] += d0 * X[i][d];
          J[1][d] += d1 * X[i][d];
          J[2][d] += d2 * X[i][d];
        }
      }
      const double dJ = det3v(J);
      if (!(dJ > 0.0)) ok = false;
      double sN = 0.0;
      for (int i = 0; i < 8; ++i) sN += N[i];
      for (int a = 0; a < 8; ++a) w[a] += rho * N[a] * sN * dJ;
    }
    return;
  }
  // CPENTA6: triangle points (2/3,1/6),(1/6,2/3),(1/6,1/6) x axial +-1/sqrt(3), weight 1/6
  const double tri[3][2] = {{2.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}, {1.0 / 6.0, 1.0 / 6.0}};
  for (int ti = 0; ti < 3; ++ti)
    for (int ai = 0; ai < 2; ++ai) {
      const double L1 = tri[ti][0], L2 = tri[ti][1], L3 = 1.0 - L1 - L2;
      const double z = ai ? kG3 : -kG3, zm = 0.5 * (1.0 - z), zp = 0.5 * (1.0 + z);
      const double N[6] = {L1 * zm, L2 * zm, L3 * zm, L1 * zp, L2 * zp, L3 * zp};
      const double dL1[6] = {zm, 0.0, -zm, zp, 0.0, -zp};
      const double dL2[6] = {0.0, zm, -zm, 0.0, zp, -zp};
      const double dz[6] = {-0.5 * L1, -0.5 * L2, -0.5 * L3, 0.5 * L1, 0.5 * L2, 0.5 * L3};
      double J[3][3] = {};
      for (int i = 0; i < 6; ++i)
        for (int d = 0; d < 3; ++d) {
          J[0][d] += dL1[i] * X[i][d];
          J[1][d] += dL2[i] * X[i][d];
          J[2][d] += dz[i] * X[i][d];
        }
      const double dJ = det3v(J);
      if (!(dJ > 0.0)) ok = false;
      double sN = 0.0;
      for (int i = 0; i < 6; ++i) sN += N[i];
      for (int a = 0; a < 6; ++a) w[a] += rho * N[a] * sN * dJ * (1.0 / 6.0);
    }
}

// rho t int(N_a) dA of a CQUAD4 (2x2 Gauss on the element's plane).
__device__ void quad_shape_integrals(const double (&X)[8][3], double rho, double t,
                                     double (&w)[8], bool& ok) {
  for (int a = 0; a < 8; ++a) w[a] = 0.0;
  double e1[3], e2[3], e3[3];
  const double v12[3] = {X[1][0] - X[0][0], X[1][1] - X[0][1], X[1][2] - X[0][2]};
  const double v14[3] = {X[3][0] - X[0][0], X[3][1] - X[0][1], X[3][2] - X[0][2]};
  e3[0] = v12[1] * v14[2] - v12[2] * v14[1];
  e3[1] = v12[2] * v14[0] - v12[0] * v14[2];
  e3[2] = v12[0] * v14[1] - v12[1] * v14[0];
  const double n3 = sqrt(e3[0] * e3[0] + e3[1] * e3[1] + e3[2] * e3[2]);
  const double n1 = sqrt(v12[0] * v12[0] + v12[1] * v12[1] + v12[2] * v12[2]);
  for (int d = 0; d < 3; ++d) {
    e3[d] /= n3;
    e1[d] = v12[d] / n1;
  }
  e2[0] = e3[1] * e1[2] - e3[2] * e1[1];
  e2[1] = e3[2] * e1[0] - e3[0] * e1[2];
  e2[2] = e3[0] * e1[1] - e3[1] * e1[0];
  double xl[4], yl[4];
  for (int n = 0; n < 4; ++n) {
    xl[n] = X[n][0] * e1[0] + X[n][1] * e1[1] + X[n][2] * e1[2];
    yl[n] = X[n][0] * e2[0] + X[n][1] * e2[1] + X[n][2] * e2[2];
  }
  const double mpa = rho * t;
  for (int gi = 0; gi < 2; ++gi)
    for (int gj = 0; gj < 2; ++gj) {
      const double xi = gi ? kG3 : -kG3, eta = gj ? kG3 : -kG3;
      const double N[4] = {0.25 * (1 - xi) * (1 - eta), 0.25 * (1 + xi) * (1 - eta),
                           0.25 * (1 + xi) * (1 + eta), 0.25 * (1 - xi) * (1 + eta)};
      const double dxi[4] = {-0.25 * (1 - eta), 0.25 * (1 - eta), 0.25 * (1 + eta), -0.25 * (1 + eta)};
      const double det[4] = {-0.25 * (1 - xi), -0.25 * (1 + xi), 0.25 * (1 + xi), 0.25 * (1 - xi)};
      double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
      for (int n = 0; n < 4; ++n) {
        J00 += dxi[n] * xl[n];
        J01 += dxi[n] * yl[n];
        J10 += det[n] * xl[n];
        J11 += det[n] * yl[n];
      }
      const double dJ = J00 * J11 - J01 * J10;
      if (!(dJ > 0.0)) ok = false;
      const double sN = N[0] + N[1] + N[2] + N[3];
      for (int a = 0; a < 4; ++a) w[a] += mpa * N[a] * sN * dJ;
    }
}

// GRAV: one element per thread, f = w_a * accel on the translational DOFs.
__global__ void k_rhs_grav(int nelem, const int* __restrict__ etype, const int* __restrict__ eptr,
                           const int* __restrict__ conn, const double* __restrict__ xyz,
                           const double* __restrict__ rho, const double* __restrict__ thick,
                           const int* __restrict__ eq, double gx, double gy, double gz,
                           double* __restrict__ F, int* __restrict__ err) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const int et = etype[e], o = eptr[e], nn = eptr[e + 1] - o;
  if (rho[e] == 0.0) return;  // zero mass matrix: skipped like the reference
  double X[8][3] = {};
  for (int a = 0; a < nn && a < 8; ++a)
    for (int d = 0; d < 3; ++d) X[a][d] = xyz[3 * conn[o + a] + d];
  double w[8];
  bool ok = true;
  if (et == VBH_CQUAD4) quad_shape_integrals(X, rho[e], thick ? thick[e] : 0.0, w, ok);
  else solid_shape_integrals(et, X, rho[e], w, ok);
  if (!ok) {
    atomicOr(&err[0], kErrJacobian);
    atomicMin(&err[1], e);
    return;
  }
  const double g[3] = {gx, gy, gz};
  double fmax = 0.0;
  for (int a = 0; a < nn; ++a)
    for (int d = 0; d < 3; ++d) fmax = fmax > fabs(w[a] * g[d]) ? fmax : fabs(w[a] * g[d]);
  if (fmax < 1e-30) return;
  for (int a = 0; a < nn; ++a)
    for (int d = 0; d < 3; ++d) {
      const int q = eq[6 * conn[o + a] + d];
      const double v = w[a] * g[d];
      if (q >= 0 && v != 0.0) atomicAdd(&F[q], v);
    }
}

__global__ void k_init_err4(int* err) {
  err[0] = 0;
  err[1] = 0x7fffffff;
}

__global__ void k_iota(int m, int* __restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) a[i] = i;
}

}  // namespace

void Ctx::assemble_rhs(const double* nodal, int nface, const int* face_nodes, const double* face_p,
                       const double* grav, const double* rho, double* F, bool dev, bool full) {
  if (!mesh.valid) throw_error(VBH_ERR_STATE, "vbh_assemble_rhs: no assembled mesh in this context");
  if (!F) throw_error(VBH_ERR_ARG, "vbh_assemble_rhs: null F");
  if (nface < 0 || (nface > 0 && (!face_nodes || !face_p)))
    throw_error(VBH_ERR_ARG, "vbh_assemble_rhs: bad face arrays");
  if (grav && !rho) throw_error(VBH_ERR_ARG, "vbh_assemble_rhs: GRAV needs element densities");
  for (int i = 0; i < 4 * nface; ++i)
    if (face_nodes[i] < 0 || face_nodes[i] >= mesh.nnodes)
      throw_error(VBH_ERR_ARG, "vbh_assemble_rhs: face node index out of range");
  select_stream_device();
  hipStream_t s = stream;
  DBuf<double> dF, dnodal, dfp, drho, dFfull;
  DBuf<int> dfn, derr, dfull;
  // full layout: every node DOF is its own equation (index 6 node + dof)
  const int nout = full ? 6 * mesh.nnodes : mesh.n;
  const int* eq = mesh.eq.p;
  if (full) {
    dfull.resize(static_cast<size_t>(nout));
    if (nout > 0) hipLaunchKernelGGL(k_iota, dim3((nout + 255) / 256), dim3(256), 0, s, nout, dfull.p);
    eq = dfull.p;
  }
  // under MPCs the loads go to the mesh numbering first, then F_red = T^T F
  const bool reduce = !full && mesh.mpc.active;
  const int nres = reduce ? mesh.mpc.n_red : nout;
  double* pF = F;
  if (!dev) {
    dF.resize(static_cast<size_t>(std::max(nres, 1)));
    pF = dF.p;
  }
  double* pOut = pF;
  if (reduce) {
    dFfull.resize(static_cast<size_t>(std::max(nout, 1)));
    pF = dFfull.p;
  }
  launch_fill_f64(nout, pF, 0.0, s);
  derr.resize(2);
  hipLaunchKernelGGL(k_init_err4, dim3(1), dim3(1), 0, s, derr.p);
  if (nodal) {
    dnodal.upload(nodal, static_cast<size_t>(6) * mesh.nnodes, s);
    hipLaunchKernelGGL(k_rhs_nodal, dim3((6 * mesh.nnodes + 255) / 256), dim3(256), 0, s,
                       mesh.nnodes, eq, dnodal.p, pF);
  }
  if (nface > 0) {
    dfn.upload(face_nodes, static_cast<size_t>(4) * nface, s);
    dfp.upload(face_p, static_cast<size_t>(4) * nface, s);
    hipLaunchKernelGGL(k_rhs_pload4, dim3((nface + 255) / 256), dim3(256), 0, s, nface, dfn.p,
                       dfp.p, mesh.xyz.p, eq, pF);
  }
  if (grav && (grav[0] != 0.0 || grav[1] != 0.0 || grav[2] != 0.0)) {
    drho.upload(rho, static_cast<size_t>(mesh.nelem), s);
    hipLaunchKernelGGL(k_rhs_grav, dim3((mesh.nelem + 255) / 256), dim3(256), 0, s, mesh.nelem,
                       mesh.et.p, mesh.ep.p, mesh.conn.p, mesh.xyz.p, drho.p,
                       mesh.t.n ? mesh.t.p : nullptr, eq, grav[0], grav[1], grav[2], pF,
                       derr.p);
  }
  if (reduce) mesh.mpc.reduce(pF, pOut, s);
  int herr[2];
  VBH_HIP(hipMemcpyAsync(herr, derr.p, sizeof herr, hipMemcpyDeviceToHost, s));
  if (!dev)
    VBH_HIP(hipMemcpyAsync(F, dF.p, sizeof(double) * nres, hipMemcpyDeviceToHost, s));
  VBH_HIP(hipStreamSynchronize(s));
  if (herr[0])
    throw_error(VBH_ERR_SOLVER,
                "element " + std::to_string(herr[1]) + ": non-positive Jacobian in mass_matrix");
}

// R[6 node + dof] = (sum_e K_e u_e) - F_applied over every node DOF: the
// support reaction at constrained DOFs, minus the equilibrium residual at free
// ones (no reference counterpart).
void Ctx::reactions(const double* u, const double* nodal, int nface, const int* face_nodes,
                    const double* face_p, const double* grav, const double* rho, double* R) {
  if (!R) throw_error(VBH_ERR_ARG, "vbh_reactions: null R");
  const size_t m = 6 * static_cast<size_t>(mesh.nnodes);
  std::vector<double> fa(m);
  element_forces(u, R, false);
  assemble_rhs(nodal, nface, face_nodes, face_p, grav, rho, fa.data(), false, true);
  for (size_t i = 0; i < m; ++i) R[i] -= fa[i];
}

}  // namespace vbh

extern "C" {

int vbh_reactions(vbh_ctx* h, const double* u, const double* nodal, int nface,
                  const int* face_nodes, const double* face_p, const double* grav,
                  const double* rho, double* R) {
  if (!h) {
    vbh::set_error("vbh_reactions: null ctx");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] { c->reactions(u, nodal, nface, face_nodes, face_p, grav, rho, R); });
}


int vbh_assemble_rhs(vbh_ctx* h, const double* nodal, int nface, const int* face_nodes,
                     const double* face_p, const double* grav, const double* rho, double* F) {
  if (!h) {
    vbh::set_error("vbh_assemble_rhs: null ctx");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] { c->assemble_rhs(nodal, nface, face_nodes, face_p, grav, rho, F, false); });
}

int vbh_assemble_rhs_device(vbh_ctx* h, const double* nodal, int nface, const int* face_nodes,
                            const double* face_p, const double* grav, const double* rho,
                            double* dF) {
  if (!h) {
    vbh::set_error("vbh_assemble_rhs_device: null ctx");
    return VBH_ERR_ARG;
  }
  auto* c = reinterpret_cast<vbh::Ctx*>(h);
  return vbh::guarded([&] { c->assemble_rhs(nodal, nface, face_nodes, face_p, grav, rho, dF, true); });
}

}  // extern "C"
