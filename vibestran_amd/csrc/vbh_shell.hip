// vibestran_amd/csrc/vbh_shell.hip -- batched CQUAD4 shell stiffness on gfx950.
//
// One wave64 per element (kElemWaves elements per block), replacing the
// reference's per-element host call CQuad4::stiffness_matrix
// (src/elements/cquad4.cpp:580-670):
//   a. every lane builds the shell frame (compute_shell_frame, 131-156): e1 along
//      node1->node2, e3 = e1 x (node1->node4) normalised, e2 = e3 x e1, and the
//      in-plane coordinates xl/yl as dot products of the absolute positions;
//   b. lanes fill the strain-displacement columns of the four 2x2 Gauss points
//      into LDS: MITC4+ tied membrane strains (311-415) or the standard
//      membrane B (Mindlin), MITC4 tied transverse shear (417-489) or the
//      Mindlin shear B, and the bending curvature gradients (254-285);
//   c. lanes form the nonzero local blocks, membrane 8x8 and plate 12x12 over
//      (rx, ry, w) per node, with the PSHELL section A = t*Dp, D = (12I/t^3)
//      t^3/12 Dp, transverse shear ts/t*t*G (build_shell_section 189-218); the
//      Mindlin form moves the in-plane shear term to a one-point rule at the
//      centre (647-655);
//   d. the drilling penalty 1e-6 x the largest rotational diagonal
//      (compute_drilling_stiffness 109-129) is added, and each lane forms its
//      global entries of T^T Kl T, T = blockdiag(R, [-R1; R0; R2]) per node,
//      before the atomic scatter into the lower CSR with the exact-zero drop
//      of SparseMatrixBuilder::add_element_stiffness (sparse_matrix.cpp:30-33).
#include <algorithm>

#include "vbh_internal.hpp"
#include "vbh_elem.hpp"

namespace vbh {
namespace {

__constant__ PairTable<24> c_spairs;

constexpr double kGp = 0.57735026918962584;  // 1/sqrt(3)

struct Frame {
  double R[3][3];  // rows e1, e2, e3
  double xl[4], yl[4];
};

__device__ __forceinline__ void make_frame(const double (&X)[4][3], Frame& f) {
  const double v12[3] = {X[1][0] - X[0][0], X[1][1] - X[0][1], X[1][2] - X[0][2]};
  const double v14[3] = {X[3][0] - X[0][0], X[3][1] - X[0][1], X[3][2] - X[0][2]};
  double e3[3] = {v12[1] * v14[2] - v12[2] * v14[1], v12[2] * v14[0] - v12[0] * v14[2],
                  v12[0] * v14[1] - v12[1] * v14[0]};
  const double n3 = sqrt(e3[0] * e3[0] + e3[1] * e3[1] + e3[2] * e3[2]);
  const double n1 = sqrt(v12[0] * v12[0] + v12[1] * v12[1] + v12[2] * v12[2]);
  double e1[3];
  for (int d = 0; d < 3; ++d) {
    e3[d] /= n3;
    e1[d] = v12[d] / n1;
  }
  const double e2[3] = {e3[1] * e1[2] - e3[2] * e1[1], e3[2] * e1[0] - e3[0] * e1[2],
                        e3[0] * e1[1] - e3[1] * e1[0]};
  for (int d = 0; d < 3; ++d) {
    f.R[0][d] = e1[d];
    f.R[1][d] = e2[d];
    f.R[2][d] = e3[d];
  }
  for (int n = 0; n < 4; ++n) {
    f.xl[n] = X[n][0] * e1[0] + X[n][1] * e1[1] + X[n][2] * e1[2];
    f.yl[n] = X[n][0] * e2[0] + X[n][1] * e2[1] + X[n][2] * e2[2];
  }
}

// a[n] for a run-time n in 0..3 / 0..1 as selects: private arrays indexed at
// run time live in scratch memory, selected ones stay in registers.
__device__ __forceinline__ double sel4(const double (&a)[4], int n) {
  return n == 0 ? a[0] : (n == 1 ? a[1] : (n == 2 ? a[2] : a[3]));
}
__device__ __forceinline__ double sel2(const double (&a)[2], int c) { return c == 0 ? a[0] : a[1]; }

// Bilinear shape functions and their parametric derivatives.
struct Shape {
  double N[4], dxi[4], deta[4];
};
__device__ __forceinline__ Shape shape(double xi, double eta) {
  Shape s;
  s.N[0] = 0.25 * (1 - xi) * (1 - eta);
  s.N[1] = 0.25 * (1 + xi) * (1 - eta);
  s.N[2] = 0.25 * (1 + xi) * (1 + eta);
  s.N[3] = 0.25 * (1 - xi) * (1 + eta);
  s.dxi[0] = -0.25 * (1 - eta);
  s.dxi[1] = 0.25 * (1 - eta);
  s.dxi[2] = 0.25 * (1 + eta);
  s.dxi[3] = -0.25 * (1 + eta);
  s.deta[0] = -0.25 * (1 - xi);
  s.deta[1] = -0.25 * (1 + xi);
  s.deta[2] = 0.25 * (1 + xi);
  s.deta[3] = 0.25 * (1 - xi);
  return s;
}

// Point data of evaluate_quad_point (cquad4.cpp:220-252): J rows are
// d(x,y)/dxi and d(x,y)/deta.
struct Point {
  Shape s;
  double J[2][2], Ji[2][2], det, dN[2][4];
};
__device__ __forceinline__ void point(const Frame& f, double xi, double eta, Point& q) {
  q.s = shape(xi, eta);
  q.J[0][0] = q.J[0][1] = q.J[1][0] = q.J[1][1] = 0.0;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    q.J[0][0] += q.s.dxi[n] * f.xl[n];
    q.J[0][1] += q.s.dxi[n] * f.yl[n];
    q.J[1][0] += q.s.deta[n] * f.xl[n];
    q.J[1][1] += q.s.deta[n] * f.yl[n];
  }
  q.det = q.J[0][0] * q.J[1][1] - q.J[0][1] * q.J[1][0];
  const double id = 1.0 / (q.det > 0.0 ? q.det : 1.0);
  q.Ji[0][0] = q.J[1][1] * id;
  q.Ji[0][1] = -q.J[0][1] * id;
  q.Ji[1][0] = -q.J[1][0] * id;
  q.Ji[1][1] = q.J[0][0] * id;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    q.dN[0][n] = q.Ji[0][0] * q.s.dxi[n] + q.Ji[0][1] * q.s.deta[n];
    q.dN[1][n] = q.Ji[1][0] * q.s.dxi[n] + q.Ji[1][1] * q.s.deta[n];
  }
}

// MITC4+ membrane interpolation coefficients (cquad4.cpp:311-370).
struct MitcCoef {
  double top, bottom, right, left, center;
};
__device__ __forceinline__ int mitc_coefficients(const Frame& f, MitcCoef& a) {
  int bad = 0;
  const double tie[5][2] = {{0, 1}, {0, -1}, {1, 0}, {-1, 0}, {0, 0}};
#pragma unroll
  for (int p = 0; p < 5; ++p) {  // tying points must map with positive Jacobian
    Point q;
    point(f, tie[p][0], tie[p][1], q);
    if (!(q.det > 0.0)) bad |= kErrJacobian;
  }
  const double RS[4] = {-1, 1, 1, -1}, SS[4] = {-1, -1, 1, 1};
  double xr[2] = {0, 0}, xs[2] = {0, 0}, xd[2] = {0, 0};
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const double p[2] = {f.xl[n], f.yl[n]};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      xr[k] += 0.25 * RS[n] * p[k];
      xs[k] += 0.25 * SS[n] * p[k];
      xd[k] += 0.25 * (RS[n] * SS[n]) * p[k];
    }
  }
  const double det = xr[0] * xs[1] - xs[0] * xr[1];
  const double rel = sqrt(xr[0] * xr[0] + xr[1] * xr[1]) * sqrt(xs[0] * xs[0] + xs[1] * xs[1]);
  if (rel < 1e-30 || fabs(det) < 1e-6 * rel) return bad | kErrMitcGeometry;
  const double mr[2] = {xs[1] / det, -xs[0] / det};
  const double ms[2] = {-xr[1] / det, xr[0] / det};
  const double cr = xd[0] * mr[0] + xd[1] * mr[1];
  const double cs = xd[0] * ms[0] + xd[1] * ms[1];
  const double dd = cr * cr + cs * cs - 1.0;
  if (fabs(dd) < 1e-12) return bad | kErrMitcDistort;
  a.top = cr * (cr - 1.0) / (2.0 * dd);
  a.bottom = cr * (cr + 1.0) / (2.0 * dd);
  a.right = cs * (cs - 1.0) / (2.0 * dd);
  a.left = cs * (cs + 1.0) / (2.0 * dd);
  a.center = 2.0 * cr * cs / dd;
  return bad;
}

// Covariant membrane strain rows of column k at a tying point
// (direct_covariant_membrane_rows): returns e_rr, e_ss, e_rs coefficients.
__device__ __forceinline__ void cov_membrane(const Frame& f, double xi, double eta, int k,
                                             double& rr, double& ss, double& rs) {
  Point q;
  point(f, xi, eta, q);
  const int n = k >> 1, c = k & 1;
  const double gr = sel2(q.J[0], c), gs = sel2(q.J[1], c);
  const double dxi = sel4(q.s.dxi, n), deta = sel4(q.s.deta, n);
  rr = gr * dxi;
  ss = gs * deta;
  rs = 0.5 * (gr * deta + gs * dxi);
}

// Membrane B column k (u or v of node k/2) at Gauss point (r, s).
__device__ __forceinline__ void membrane_col(const Frame& f, const MitcCoef& a, int form, double r, double s,
                             int k, double out[3]) {
  Point q;
  point(f, r, s, q);
  const int n = k >> 1, c = k & 1;
  if (form == 1) {
    const double d0 = sel4(q.dN[0], n), d1 = sel4(q.dN[1], n);
    out[0] = c == 0 ? d0 : 0.0;
    out[1] = c == 1 ? d1 : 0.0;
    out[2] = c == 0 ? d1 : d0;
    return;
  }
  double t_rr, t_ss, t_rs, b_rr, b_ss, b_rs, r_rr, r_ss, r_rs, l_rr, l_ss, l_rs, c_rr, c_ss, c_rs;
  cov_membrane(f, 0.0, 1.0, k, t_rr, t_ss, t_rs);
  cov_membrane(f, 0.0, -1.0, k, b_rr, b_ss, b_rs);
  cov_membrane(f, 1.0, 0.0, k, r_rr, r_ss, r_rs);
  cov_membrane(f, -1.0, 0.0, k, l_rr, l_ss, l_rs);
  cov_membrane(f, 0.0, 0.0, k, c_rr, c_ss, c_rs);
  const double err = 0.5 * (1.0 - 2.0 * a.top + s + 2.0 * a.top * s * s) * t_rr +
                     0.5 * (1.0 - 2.0 * a.bottom - s + 2.0 * a.bottom * s * s) * b_rr +
                     a.right * (-1.0 + s * s) * r_ss + a.left * (-1.0 + s * s) * l_ss +
                     a.center * (-1.0 + s * s) * c_rs;
  const double ess = a.top * (-1.0 + r * r) * t_rr + a.bottom * (-1.0 + r * r) * b_rr +
                     0.5 * (1.0 - 2.0 * a.right + r + 2.0 * a.right * r * r) * r_ss +
                     0.5 * (1.0 - 2.0 * a.left - r + 2.0 * a.left * r * r) * l_ss +
                     a.center * (-1.0 + r * r) * c_rs;
  const double ers = 0.25 * (r + 4.0 * a.top * r * s) * t_rr +
                     0.25 * (-r + 4.0 * a.bottom * r * s) * b_rr +
                     0.25 * (s + 4.0 * a.right * r * s) * r_ss +
                     0.25 * (-s + 4.0 * a.left * r * s) * l_ss + (1.0 + a.center * r * s) * c_rs;
  const double A = q.Ji[0][0], B = q.Ji[1][0], C = q.Ji[0][1], D = q.Ji[1][1];
  out[0] = A * A * err + C * C * ess + 2.0 * A * C * ers;
  out[1] = B * B * err + D * D * ess + 2.0 * B * D * ers;
  out[2] = 2.0 * A * B * err + 2.0 * C * D * ess + 2.0 * (A * D + B * C) * ers;
}

// Covariant transverse shear of plate column k = 3n + c (c: 0 rx, 1 ry, 2 w)
// at a tying point: (gamma_xi, gamma_eta) rows.
__device__ __forceinline__ void cov_shear(const Frame& f, double xi, double eta, int k,
                                          double& gxi, double& geta) {
  Point q;
  point(f, xi, eta, q);
  const int n = k / 3, c = k % 3;
  if (c == 2) {
    gxi = sel4(q.s.dxi, n);
    geta = sel4(q.s.deta, n);
  } else {
    const double Nn = sel4(q.s.N, n);
    gxi = -Nn * sel2(q.J[0], c);
    geta = -Nn * sel2(q.J[1], c);
  }
}

// Transverse shear B column k at Gauss point (r, s): MITC4 tying or Mindlin.
__device__ __forceinline__ void shear_col(const Frame& f, int form, double r, double s, int k, double out[2]) {
  Point q;
  point(f, r, s, q);
  const int n = k / 3, c = k % 3;
  if (form == 1) {
    const double d0 = sel4(q.dN[0], n), d1 = sel4(q.dN[1], n), Nn = sel4(q.s.N, n);
    out[0] = c == 2 ? d0 : (c == 0 ? -Nn : 0.0);
    out[1] = c == 2 ? d1 : (c == 1 ? -Nn : 0.0);
    return;
  }
  double xb, eb, xt, et, xr, er, xl, el, xc, ec;
  cov_shear(f, 0.0, -1.0, k, xb, eb);
  cov_shear(f, 0.0, 1.0, k, xt, et);
  cov_shear(f, 1.0, 0.0, k, xr, er);
  cov_shear(f, -1.0, 0.0, k, xl, el);
  cov_shear(f, 0.0, 0.0, k, xc, ec);
  const double cxi = 0.5 * (1.0 - s) * xb + 0.5 * (1.0 + s) * xt + r * (xc - 0.5 * (xb + xt));
  const double ceta = 0.5 * (1.0 - r) * el + 0.5 * (1.0 + r) * er + s * (ec - 0.5 * (el + er));
  out[0] = q.Ji[0][0] * cxi + q.Ji[0][1] * ceta;
  out[1] = q.Ji[1][0] * cxi + q.Ji[1][1] * ceta;
}

struct ShellLds {
  double Bm[4][8][3];   // membrane B columns per Gauss point
  double Bs[4][12][2];  // transverse shear B columns
  double dN[4][2][4];   // physical shape gradients (bending)
  double w[4];          // Gauss weights (detJ)
  double c2[8];         // Mindlin centre in-plane shear row
  double w0;
  double Km[8][8];      // local membrane block (both triangles)
  double Kp[12][12];    // local plate block, columns (rx, ry, w) per node
  double drill;
  double dg[20];  // Kl diagonal before differentiation (thickness mode)
};

// One entry of the B^T D B product of a B with M strain rows, in the
// reference's operation order (Eigen-free restatement of the element's
// K += B^T (D B) w, cquad4.cpp:622-655): (sum_k B[k][i] (sum_l D[k][l] B[l][j])) w,
// every sum from 0.0 in increasing index order. With floating-point
// contraction off (this file is compiled with -ffp-contract=off) the entry is
// bit-identical to the CPU oracle's, so entries that cancel to exactly 0.0
// there do so here, and the CSR pattern's exact-zero drop matches.
template <int M, typename BCOL>
__device__ __forceinline__ double btdb_entry(const double (&D)[M][M], BCOL col, int i, int j,
                                             double w) {
  double bi[M], bj[M];
  col(i, bi);
  col(j, bj);
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    double db = 0.0;
#pragma unroll
    for (int l = 0; l < M; ++l) db += D[k][l] * bj[l];
    s += bi[k] * db;
  }
  return s * w;
}

// Bending curvature column of plate column k: (kx, ky, kxy).
__device__ __forceinline__ void bend_col(const ShellLds& L, int g, int k, double b[3]) {
  const int n = k / 3, c = k % 3;
  const double dx = L.dN[g][0][n], dy = L.dN[g][1][n];
  b[0] = c == 0 ? dx : 0.0;
  b[1] = c == 1 ? dy : 0.0;
  b[2] = c == 0 ? dy : (c == 1 ? dx : 0.0);
}

// x^T Dp y for the plane-stress pattern [d00 d01 0; d01 d00 0; 0 0 d22].
__device__ __forceinline__ double dform(double d00, double d01, double d22, const double* x,
                                        const double* y) {
  return d00 * (x[0] * y[0] + x[1] * y[1]) + d01 * (x[0] * y[1] + x[1] * y[0]) + d22 * x[2] * y[2];
}

// Local stiffness entry Kl[6i+p][6j+q] from the LDS blocks.
__device__ __forceinline__ double klocal(const ShellLds& L, int i, int p, int j, int q) {
  if (p < 2 && q < 2) return L.Km[2 * i + p][2 * j + q];
  if (p >= 2 && p < 5 && q >= 2 && q < 5) {
    const int pp = (p == 2) ? 2 : p - 3, qq = (q == 2) ? 2 : q - 3;
    return L.Kp[3 * i + pp][3 * j + qq];
  }
  if (p == 5 && q == 5 && i == j) return L.drill;
  return 0.0;
}

// Column of T for global DOF a (0..5) of a node: local DOF index of each term
// and its coefficient; T = blockdiag(R, [-R1; R0; R2]).
__device__ __forceinline__ double rsel(const Frame& f, int c, int b) {
  return b == 0 ? f.R[c][0] : (b == 1 ? f.R[c][1] : f.R[c][2]);
}

__device__ __forceinline__ void tcol(const Frame& f, int a, double t[3], int l[3]) {
  if (a < 3) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      t[c] = rsel(f, c, a);
      l[c] = c;
    }
  } else {
    const int b = a - 3;
    t[0] = -rsel(f, 1, b);
    t[1] = rsel(f, 0, b);
    t[2] = rsel(f, 2, b);
    l[0] = 3;
    l[1] = 4;
    l[2] = 5;
  }
}

// MODE 3: out[6 node + dof] += (K_e u_e) in the basic frame.
// MODE 0: assemble (scatter into the lower CSR).
// MODE 1: out[e] = -u_e^T (dK_e/dE_e) v_e = -u_e^T K_e v_e / E_e (K_e is linear in E).
// MODE 2: out[e] = -u_e^T (dK_e/dt_e) v_e: membrane and transverse shear scale
//         with t, bending with t^3, the drilling penalty with the rotational
//         diagonal entry it was taken from.
// 2 resident waves per SIMD asked of the compiler (256 VGPRs, a few spilled):
// 1 wave without the bound (256 VGPRs + AGPRs). Tube assembly 8.8 -> 6.9 ms,
// adjoint contractions 12.5 -> 10.4 ms, element forces 12.1 -> 10.1 ms; a
// bound of 4 waves spills too much (assembly 10.9 ms).
template <int MODE>
__global__ __launch_bounds__(64 * kElemWaves, 2) void k_quad4(
    int nsel, const int* __restrict__ sel, const int* __restrict__ eptr,
    const int* __restrict__ conn, const double* __restrict__ xyz, const double* __restrict__ Ev,
    const double* __restrict__ nuv, const double* __restrict__ thick, int form,
    const int* __restrict__ gd, const int* __restrict__ rp, const int* __restrict__ ci,
    double* __restrict__ val, int* __restrict__ touched, int* __restrict__ err,
    const double* __restrict__ uvec, const double* __restrict__ vvec, double* __restrict__ out) {
  __shared__ ShellLds lds_all[kElemWaves];
  __shared__ double facc_all[kElemWaves][MODE == 3 ? 24 : 1];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  ShellLds& L = lds_all[wave];
  double* facc = facc_all[wave];
  for (int base = blockIdx.x * kElemWaves; base < nsel; base += gridDim.x * kElemWaves) {
    const int s = base + wave;
    const bool active = s < nsel;
    const int e = active ? sel[s] : sel[base];
    double X[4][3];
    const int o = eptr[e];
    for (int a = 0; a < 4; ++a)
      for (int d = 0; d < 3; ++d) X[a][d] = xyz[3 * conn[o + a] + d];
    Frame f;
    make_frame(X, f);
    MitcCoef mc{0, 0, 0, 0, 0};
    int bad = 0;
    if (form == 0) bad |= mitc_coefficients(f, mc);
    const double E = Ev[e], nu = nuv[e], t = thick[e];
    const double c = E / (1.0 - nu * nu);
    const double G = E / (2.0 * (1.0 + nu));
    const double ib = 1.0 * t * t * t / 12.0;  // PSHELL 12I/T^3 = 1.0
    const double ts = 0.833333 * t * G;         // PSHELL TS/T = 0.833333
    const double A00 = t * c, A01 = t * c * nu, A22 = t * c * (1.0 - nu) / 2.0;
    const double D00 = ib * c, D01 = ib * c * nu, D22 = ib * c * (1.0 - nu) / 2.0;
    // section matrices as the reference forms them (cquad4.cpp:189-218):
    // Dp plane stress, membrane A = t Dp, bending ib Dp, shear tst t G I2
    const double Dp[3][3] = {{c, c * nu, 0.0}, {c * nu, c, 0.0}, {0.0, 0.0, c * (1.0 - nu) / 2.0}};
    double Am[3][3], Db[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        Am[a][b] = t * Dp[a][b];
        Db[a][b] = ib * Dp[a][b];
      }
    const double A22x = Am[2][2];  // Mindlin: in-plane shear at the centre only
    if (form == 1) Am[2][2] = 0.0;
    const double Ts[2][2] = {{ts, 0.0}, {0.0, ts}};
    const double Ash[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, A22x}};

    // b. strain-displacement columns: 32 membrane, 48 shear, 4 points, centre
    for (int it = lane; it < 85; it += 64) {
      if (it < 32) {
        const int g = it >> 3, k = it & 7;
        const double r = (g & 2) ? kGp : -kGp, sv = (g & 1) ? kGp : -kGp;
        double b[3];
        membrane_col(f, mc, form, r, sv, k, b);
        for (int m = 0; m < 3; ++m) L.Bm[g][k][m] = b[m];
      } else if (it < 80) {
        const int g = (it - 32) / 12, k = (it - 32) % 12;
        const double r = (g & 2) ? kGp : -kGp, sv = (g & 1) ? kGp : -kGp;
        double b[2];
        shear_col(f, form, r, sv, k, b);
        L.Bs[g][k][0] = b[0];
        L.Bs[g][k][1] = b[1];
      } else if (it < 84) {
        const int g = it - 80;
        Point q;
        point(f, (g & 2) ? kGp : -kGp, (g & 1) ? kGp : -kGp, q);
        if (!(q.det > 0.0)) bad |= kErrJacobian;
        for (int n = 0; n < 4; ++n) {
          L.dN[g][0][n] = q.dN[0][n];
          L.dN[g][1][n] = q.dN[1][n];
        }
        L.w[g] = q.det;
      } else {
        Point q;
        point(f, 0.0, 0.0, q);
        for (int n = 0; n < 4; ++n) {
          L.c2[2 * n] = q.dN[1][n];
          L.c2[2 * n + 1] = q.dN[0][n];
        }
        L.w0 = q.det * 4.0;
      }
    }
    __syncthreads();

    // c. local membrane (8 x 8) and plate (12 x 12) blocks, every entry in
    //    the reference's order: per Gauss point the membrane B^T A B w, then
    //    the bending B^T Db B w and the shear B^T Ts B w (cquad4.cpp:622-655),
    //    the Mindlin centre in-plane shear last
    if (MODE != 2) {
      for (int it = lane; it < 64 + 144; it += 64) {
        if (it < 64) {
          const int i = it >> 3, j = it & 7;
          double acc = 0.0;
          for (int g = 0; g < 4; ++g) {
            auto col = [&](int k, double (&b)[3]) {
              for (int m = 0; m < 3; ++m) b[m] = L.Bm[g][k][m];
            };
            acc += btdb_entry<3>(Am, col, i, j, L.w[g]);
          }
          if (form == 1) {
            auto col0 = [&](int k, double (&b)[3]) {  // standard membrane B at the centre
              b[0] = (k & 1) ? 0.0 : L.c2[k + 1];
              b[1] = (k & 1) ? L.c2[k - 1] : 0.0;
              b[2] = L.c2[k];
            };
            acc += btdb_entry<3>(Ash, col0, i, j, L.w0);
          }
          L.Km[i][j] = acc;
        } else {
          const int i = (it - 64) / 12, j = (it - 64) % 12;
          double acc = 0.0;
          for (int g = 0; g < 4; ++g) {
            auto bcol = [&](int k, double (&b)[3]) { bend_col(L, g, k, b); };
            auto scol = [&](int k, double (&b)[2]) {
              b[0] = L.Bs[g][k][0];
              b[1] = L.Bs[g][k][1];
            };
            acc += btdb_entry<3>(Db, bcol, i, j, L.w[g]);
            acc += btdb_entry<2>(Ts, scol, i, j, L.w[g]);
          }
          L.Kp[i][j] = acc;
        }
      }
    }
    for (int it = lane; MODE == 2 && it < 36 + 78; it += 64) {
      if (it < 36) {
        const unsigned short pr = c_spairs.v[it];
        const int i = pr >> 8, j = pr & 0xff;
        double sum = 0.0;
        for (int g = 0; g < 4; ++g)
          sum += L.w[g] * dform(A00, A01, form == 0 ? A22 : 0.0, L.Bm[g][i], L.Bm[g][j]);
        if (form == 1) sum += L.w0 * A22 * L.c2[i] * L.c2[j];
        if (MODE == 2 && i == j) L.dg[i] = sum;
        if (MODE == 2) sum /= t;  // membrane stiffness is proportional to t
        L.Km[i][j] = sum;
        L.Km[j][i] = sum;
      } else {
        const unsigned short pr = c_spairs.v[it - 36];
        const int i = pr >> 8, j = pr & 0xff;
        double sum = 0.0, dsum = 0.0;
        for (int g = 0; g < 4; ++g) {
          double bi[3], bj[3];
          bend_col(L, g, i, bi);
          bend_col(L, g, j, bj);
          const double kb = dform(D00, D01, D22, bi, bj);
          const double ks = ts * (L.Bs[g][i][0] * L.Bs[g][j][0] + L.Bs[g][i][1] * L.Bs[g][j][1]);
          sum += L.w[g] * (kb + ks);
          dsum += L.w[g] * (3.0 * kb + ks);  // t^3 bending, t shear
        }
        if (MODE == 2) {
          if (i == j) L.dg[8 + i] = sum;
          sum = dsum / t;
        }
        L.Kp[i][j] = sum;
        L.Kp[j][i] = sum;
      }
    }
    __syncthreads();

    // d. drilling penalty (cquad4.cpp:655-667 with 109-129)
    if (lane == 0) {
      // diagonal of Kl before any derivative: membrane 0..7, plate 8..19
      double dg[20];
      for (int i = 0; i < 8; ++i) dg[i] = (MODE == 2) ? L.dg[i] : L.Km[i][i];
      for (int i = 0; i < 12; ++i) dg[8 + i] = (MODE == 2) ? L.dg[8 + i] : L.Kp[i][i];
      double maxd = 0.0, maxrot = 0.0;
      int imax = 0, irot = -1;
      for (int i = 0; i < 20; ++i)
        if (fabs(dg[i]) > maxd) {
          maxd = fabs(dg[i]);
          imax = i;
        }
      for (int n = 0; n < 4; ++n)
        for (int c = 0; c < 2; ++c)
          if (fabs(dg[8 + 3 * n + c]) > maxrot) {
            maxrot = fabs(dg[8 + 3 * n + c]);
            irot = 8 + 3 * n + c;
          }
      if (maxd < 1e-10) bad |= kErrShellDiag;
      if (MODE != 2) {
        L.drill = (maxrot > 0.0) ? 1e-6 * maxrot : 1e-12 * maxd;
      } else {
        // derivative of the penalty along the entry that defines it
        const int k = (maxrot > 0.0) ? irot : imax;
        const double dk = (k < 8) ? L.Km[k][k] : L.Kp[k - 8][k - 8];
        const double sg = (L.dg[k] < 0.0) ? -1.0 : 1.0;  // d|x| = sign(x) dx
        L.drill = ((maxrot > 0.0) ? 1e-6 : 1e-12) * sg * dk;
      }
    }
    __syncthreads();

    if (MODE == 3) {
      // f_e = K_e u_e in the basic frame, summed in LDS, added to out[6 node + dof]
      const int* g = gd + static_cast<size_t>(e) * kGdStride;
      if (lane < 24) facc[lane] = 0.0;
      __syncthreads();
      for (int it = lane; it < 300; it += 64) {
        const unsigned short pr = c_spairs.v[it];
        const int I = pr >> 8, J = pr & 0xff;
        const int ni = I / 6, a = I % 6, nj = J / 6, b = J % 6;
        double ta[3], tb[3];
        int la[3], lb[3];
        tcol(f, a, ta, la);
        tcol(f, b, tb, lb);
        double kv = 0.0;
        for (int p = 0; p < 3; ++p) {
          double inner = 0.0;
          for (int q = 0; q < 3; ++q) inner += klocal(L, ni, la[p], nj, lb[q]) * tb[q];
          kv += ta[p] * inner;
        }
        const double uI = g[I] >= 0 ? uvec[g[I]] : 0.0, uJ = g[J] >= 0 ? uvec[g[J]] : 0.0;
        atomicAdd(&facc[I], kv * uJ);
        if (I != J) atomicAdd(&facc[J], kv * uI);
      }
      __syncthreads();
      if (active && lane < 24) atomicAdd(&out[6 * conn[o + lane / 6] + lane % 6], facc[lane]);
      if (active && lane == 0 && bad) {
        atomicOr(&err[0], bad);
        atomicMin(&err[1], e);
      }
      __syncthreads();
    }

    if (MODE == 1 || MODE == 2) {
      // contraction with the element's parts of u and v
      const int* g = gd + static_cast<size_t>(e) * kGdStride;
      double part = 0.0;
      for (int it = lane; it < 300; it += 64) {
        const unsigned short pr = c_spairs.v[it];
        const int I = pr >> 8, J = pr & 0xff;
        const int gi = g[I], gj = g[J];
        if (gi < 0 || gj < 0) continue;
        const int ni = I / 6, a = I % 6, nj = J / 6, b = J % 6;
        double ta[3], tb[3];
        int la[3], lb[3];
        tcol(f, a, ta, la);
        tcol(f, b, tb, lb);
        double kv = 0.0;
        for (int p = 0; p < 3; ++p) {
          double inner = 0.0;
          for (int q = 0; q < 3; ++q) inner += klocal(L, ni, la[p], nj, lb[q]) * tb[q];
          kv += ta[p] * inner;
        }
        double w = uvec[gi] * vvec[gj];
        if (I != J) w += uvec[gj] * vvec[gi];
        part += kv * w;
      }
      for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off, 64);
      if (active && lane == 0) {
        out[e] = (MODE == 1) ? -part / E : -part;
        if (bad) {
          atomicOr(&err[0], bad);
          atomicMin(&err[1], e);
        }
      }
    }

    if (MODE == 0 && active) {
      if (bad) {
        atomicOr(&err[0], bad);
        atomicMin(&err[1], e);
      }
      // entry `it` of element s at val[s * 300 + it] (its pair position; the
      // slot sums read only pairs whose DOFs are both free)
      const int* g = gd + static_cast<size_t>(e) * kGdStride;
      double* dst = val + static_cast<size_t>(s) * 300;
      for (int it = lane; it < 300; it += 64) {
        const unsigned short pr = c_spairs.v[it];
        int I = pr >> 8, J = pr & 0xff;
        const int gi = g[I], gj = g[J];
        if (gi < 0 || gj < 0) continue;
        if (gi < gj) {  // row = larger global index
          const int tmp = I;
          I = J;
          J = tmp;
        }
        const int ni = I / 6, a = I % 6, nj = J / 6, b = J % 6;
        double ta[3], tb[3];
        int la[3], lb[3];
        tcol(f, a, ta, la);
        tcol(f, b, tb, lb);
        double v = 0.0;
        for (int p = 0; p < 3; ++p) {
          double inner = 0.0;
          for (int q = 0; q < 3; ++q) inner += klocal(L, ni, la[p], nj, lb[q]) * tb[q];
          v += ta[p] * inner;
        }
        dst[it] = v;
      }
    }
    __syncthreads();
  }
}

// Covariant membrane strains (e_rr, e_ss, e_rs) of the membrane DOFs um at a
// tying point: the contraction of cov_membrane's columns with um.
__device__ void cov_membrane_strain(const Frame& f, double xi, double eta, const double* um,
                                    double& rr, double& ss, double& rs) {
  Point q;
  point(f, xi, eta, q);
  rr = ss = rs = 0.0;
  for (int k = 0; k < 8; ++k) {
    const int n = k >> 1, c = k & 1;
    const double gr = q.J[0][c], gs = q.J[1][c];
    rr += gr * q.s.dxi[n] * um[k];
    ss += gs * q.s.deta[n] * um[k];
    rs += 0.5 * (gr * q.s.deta[n] + gs * q.s.dxi[n]) * um[k];
  }
}

// Stress recovery of recover_shell_response (cquad4.cpp:758-801), one thread
// per element: the displacements go to the local frame (T u,
// to_local_displacements 491-502), the membrane strain is the MITC4+ tied
// strain (the tying-point strains contracted once, then interpolated as in
// mitc4plus_membrane_B) or the standard one (Mindlin), the curvature the
// bending B of the plate DOFs (rx, ry, w). Rows are
// [sx, sy, sxy, 0, 0, 0, mx, my, mxy, von_mises] at the centroid and at the
// four corners (recover_results, linear_static.cpp:712-753); no thermal strain.
__global__ __launch_bounds__(256) void k_quad4_recover(
    int nsel, const int* __restrict__ sel, const int* __restrict__ eptr,
    const int* __restrict__ conn, const double* __restrict__ xyz, const double* __restrict__ E,
    const double* __restrict__ nu, const double* __restrict__ thick, int form,
    const int* __restrict__ gd, const double* __restrict__ u, double* __restrict__ cen,
    double* __restrict__ cor, int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nsel) return;
  const int e = sel[i];
  const int o = eptr[e];
  double X[4][3];
  for (int a = 0; a < 4; ++a) {
    const int node = conn[o + a];
    for (int d = 0; d < 3; ++d) X[a][d] = xyz[3 * node + d];
  }
  Frame f;
  make_frame(X, f);
  const int* g = gd + static_cast<size_t>(e) * kGdStride;
  double um[8], up[12];
  for (int n = 0; n < 4; ++n) {
    double ut[3], th[3];
    for (int d = 0; d < 3; ++d) {
      const int qt = g[6 * n + d], qr = g[6 * n + 3 + d];
      ut[d] = qt >= 0 ? u[qt] : 0.0;
      th[d] = qr >= 0 ? u[qr] : 0.0;
    }
    double loc[3], rot[3];
    for (int a = 0; a < 3; ++a) {
      loc[a] = f.R[a][0] * ut[0] + f.R[a][1] * ut[1] + f.R[a][2] * ut[2];
      rot[a] = f.R[a][0] * th[0] + f.R[a][1] * th[1] + f.R[a][2] * th[2];
    }
    um[2 * n] = loc[0];
    um[2 * n + 1] = loc[1];
    up[3 * n] = -rot[1];      // local rx = -(e2 . theta)
    up[3 * n + 1] = rot[0];   // local ry = e1 . theta
    up[3 * n + 2] = loc[2];   // w
  }
  int bad = 0;
  MitcCoef a{};
  double T_rr = 0, B_rr = 0, R_ss = 0, L_ss = 0, C_rs = 0;
  if (form == 0) {
    bad |= mitc_coefficients(f, a);
    double x1, x2;
    cov_membrane_strain(f, 0.0, 1.0, um, T_rr, x1, x2);
    cov_membrane_strain(f, 0.0, -1.0, um, B_rr, x1, x2);
    cov_membrane_strain(f, 1.0, 0.0, um, x1, R_ss, x2);
    cov_membrane_strain(f, -1.0, 0.0, um, x1, L_ss, x2);
    cov_membrane_strain(f, 0.0, 0.0, um, x1, x2, C_rs);
  }
  const double c = E[e] / (1.0 - nu[e] * nu[e]);
  const double d01 = c * nu[e], d22 = c * (1.0 - nu[e]) / 2.0;
  const double t = thick[e], ib = t * t * t / 12.0;  // 12I/t^3 = 1
  const double pts[5][2] = {{0.0, 0.0}, {-1.0, -1.0}, {1.0, -1.0}, {1.0, 1.0}, {-1.0, 1.0}};
  const int npt = cor ? 5 : 1;
  for (int p = 0; p < npt; ++p) {
    const double r = pts[p][0], s = pts[p][1];
    Point q;
    point(f, r, s, q);
    if (!(q.det > 0.0)) bad |= kErrJacobian;
    double eps[3];
    if (form == 0) {
      const double err_ = 0.5 * (1.0 - 2.0 * a.top + s + 2.0 * a.top * s * s) * T_rr +
                          0.5 * (1.0 - 2.0 * a.bottom - s + 2.0 * a.bottom * s * s) * B_rr +
                          a.right * (-1.0 + s * s) * R_ss + a.left * (-1.0 + s * s) * L_ss +
                          a.center * (-1.0 + s * s) * C_rs;
      const double ess_ = a.top * (-1.0 + r * r) * T_rr + a.bottom * (-1.0 + r * r) * B_rr +
                          0.5 * (1.0 - 2.0 * a.right + r + 2.0 * a.right * r * r) * R_ss +
                          0.5 * (1.0 - 2.0 * a.left - r + 2.0 * a.left * r * r) * L_ss +
                          a.center * (-1.0 + r * r) * C_rs;
      const double ers_ = 0.25 * (r + 4.0 * a.top * r * s) * T_rr +
                          0.25 * (-r + 4.0 * a.bottom * r * s) * B_rr +
                          0.25 * (s + 4.0 * a.right * r * s) * R_ss +
                          0.25 * (-s + 4.0 * a.left * r * s) * L_ss +
                          (1.0 + a.center * r * s) * C_rs;
      const double A = q.Ji[0][0], B = q.Ji[1][0], C = q.Ji[0][1], D = q.Ji[1][1];
      eps[0] = A * A * err_ + C * C * ess_ + 2.0 * A * C * ers_;
      eps[1] = B * B * err_ + D * D * ess_ + 2.0 * B * D * ers_;
      eps[2] = 2.0 * A * B * err_ + 2.0 * C * D * ess_ + 2.0 * (A * D + B * C) * ers_;
    } else {
      eps[0] = eps[1] = eps[2] = 0.0;
      for (int n = 0; n < 4; ++n) {
        eps[0] += q.dN[0][n] * um[2 * n];
        eps[1] += q.dN[1][n] * um[2 * n + 1];
        eps[2] += q.dN[1][n] * um[2 * n] + q.dN[0][n] * um[2 * n + 1];
      }
    }
    double kap[3] = {0.0, 0.0, 0.0};
    for (int n = 0; n < 4; ++n) {
      kap[0] += q.dN[0][n] * up[3 * n];
      kap[1] += q.dN[1][n] * up[3 * n + 1];
      kap[2] += q.dN[1][n] * up[3 * n] + q.dN[0][n] * up[3 * n + 1];
    }
    const double sx = c * eps[0] + d01 * eps[1], sy = d01 * eps[0] + c * eps[1],
                 sxy = d22 * eps[2];
    const double row[10] = {sx, sy, sxy, 0.0, 0.0, 0.0,
                            ib * (c * kap[0] + d01 * kap[1]), ib * (d01 * kap[0] + c * kap[1]),
                            ib * d22 * kap[2],
                            sqrt(sx * sx - sx * sy + sy * sy + 3.0 * sxy * sxy)};
    double* out = p == 0 ? cen + 10 * static_cast<size_t>(e)
                         : cor + 80 * static_cast<size_t>(e) + 10 * (p - 1);
    for (int k = 0; k < 10; ++k) out[k] = row[k];
  }
  if (cor)
    for (int k = 40; k < 80; ++k) cor[80 * static_cast<size_t>(e) + k] = 0.0;
  if (bad) {
    atomicOr(&err[0], bad);
    atomicMin(&err[1], e);
  }
}

}  // namespace

void launch_quad4_recover(int nsel, const int* sel, const int* eptr, const int* conn,
                          const double* xyz, const double* E, const double* nu,
                          const double* thick, int form, const int* gd, const double* u,
                          double* cen, double* cor, int* err, hipStream_t s) {
  if (nsel <= 0) return;
  hipLaunchKernelGGL(k_quad4_recover, dim3((nsel + 255) / 256), dim3(256), 0, s, nsel, sel, eptr,
                     conn, xyz, E, nu, thick, form, gd, u, cen, cor, err);
}

void launch_quad4_forces(int nsel, const int* sel, const int* eptr, const int* conn,
                         const double* xyz, const double* E, const double* nu, const double* thick,
                         int form, const int* gd, const double* u, double* fout, int* err,
                         hipStream_t s) {
  if (nsel <= 0) return;
  const int blocks = std::min((nsel + kElemWaves - 1) / kElemWaves, 65535);
  hipLaunchKernelGGL(k_quad4<3>, dim3(blocks), dim3(64 * kElemWaves), 0, s, nsel, sel, eptr, conn,
                     xyz, E, nu, thick, form, gd, nullptr, nullptr, nullptr, nullptr, err, u,
                     nullptr, fout);
}

void launch_quad4_stiffness(int nsel, const int* sel, const int* eptr, const int* conn,
                            const double* xyz, const double* E, const double* nu,
                            const double* thick, int form, const int* gd, double* ke, int* err,
                            hipStream_t s) {
  if (nsel <= 0) return;
  const int blocks = std::min((nsel + kElemWaves - 1) / kElemWaves, 65535);
  hipLaunchKernelGGL(k_quad4<0>, dim3(blocks), dim3(64 * kElemWaves), 0, s, nsel, sel, eptr, conn,
                     xyz, E, nu, thick, form, gd, nullptr, nullptr, ke, nullptr, err, nullptr,
                     nullptr, nullptr);
}

void launch_quad4_sensitivity(int nsel, const int* sel, const int* eptr, const int* conn,
                              const double* xyz, const double* E, const double* nu,
                              const double* thick, int form, const int* gd, int param,
                              const double* u, const double* v, double* out, int* err,
                              hipStream_t s) {
  if (nsel <= 0) return;
  const int blocks = std::min((nsel + kElemWaves - 1) / kElemWaves, 65535);
  if (param == 0)
    hipLaunchKernelGGL(k_quad4<1>, dim3(blocks), dim3(64 * kElemWaves), 0, s, nsel, sel, eptr,
                       conn, xyz, E, nu, thick, form, gd, nullptr, nullptr, nullptr, nullptr, err,
                       u, v, out);
  else
    hipLaunchKernelGGL(k_quad4<2>, dim3(blocks), dim3(64 * kElemWaves), 0, s, nsel, sel, eptr,
                       conn, xyz, E, nu, thick, form, gd, nullptr, nullptr, nullptr, nullptr, err,
                       u, v, out);
}

}  // namespace vbh
