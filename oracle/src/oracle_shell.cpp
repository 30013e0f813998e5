// oracle/src/oracle_shell.cpp -- TEST INFRASTRUCTURE ONLY.
// CPU restatement of the reference's CQUAD4 shell stiffness
// (src/elements/cquad4.cpp); see oracle.hpp for the cited line ranges.
// Pinned by tests/test_oracle_shell.py: symmetry, the six rigid-body modes
// (translations and rotations, including drilling), membrane/bending patch
// states, orientation invariance of the spectrum, and the reference e2e decks
// cquad4_axial_force (exact), cantilever and closed-tube bending (+-8%).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "oracle.hpp"

namespace {

struct V3 {
  double x, y, z;
};
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 normalized(V3 a) {
  const double n = std::sqrt(dot(a, a));
  return {a.x / n, a.y / n, a.z / n};
}

struct QShape {
  double N[4], dxi[4], deta[4];
};
QShape qshape(double xi, double eta) {
  QShape s;
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

struct QPoint {
  double xi, eta, J[2][2], Jinv[2][2], dNdx[2][4], detJ;
  QShape s;
};

bool eval_point(const double* xl, const double* yl, double xi, double eta, QPoint& q) {
  q.xi = xi;
  q.eta = eta;
  q.s = qshape(xi, eta);
  std::memset(q.J, 0, sizeof q.J);
  for (int n = 0; n < 4; ++n) {
    q.J[0][0] += q.s.dxi[n] * xl[n];
    q.J[0][1] += q.s.dxi[n] * yl[n];
    q.J[1][0] += q.s.deta[n] * xl[n];
    q.J[1][1] += q.s.deta[n] * yl[n];
  }
  q.detJ = q.J[0][0] * q.J[1][1] - q.J[0][1] * q.J[1][0];
  if (q.detJ <= 0.0) return false;
  const double id = 1.0 / q.detJ;
  q.Jinv[0][0] = q.J[1][1] * id;
  q.Jinv[0][1] = -q.J[0][1] * id;
  q.Jinv[1][0] = -q.J[1][0] * id;
  q.Jinv[1][1] = q.J[0][0] * id;
  for (int n = 0; n < 4; ++n) {
    q.dNdx[0][n] = q.Jinv[0][0] * q.s.dxi[n] + q.Jinv[0][1] * q.s.deta[n];
    q.dNdx[1][n] = q.Jinv[1][0] * q.s.dxi[n] + q.Jinv[1][1] * q.s.deta[n];
  }
  return true;
}

// covariant membrane strain rows at a tying point (direct_covariant_membrane_rows)
struct CovRows {
  double rr[8], ss[8], rs[8];
};
bool cov_rows(const double* xl, const double* yl, double xi, double eta, CovRows& c) {
  QPoint q;
  if (!eval_point(xl, yl, xi, eta, q)) return false;
  const double grx = q.J[0][0], gry = q.J[0][1], gsx = q.J[1][0], gsy = q.J[1][1];
  for (int n = 0; n < 4; ++n) {
    c.rr[2 * n] = grx * q.s.dxi[n];
    c.rr[2 * n + 1] = gry * q.s.dxi[n];
    c.ss[2 * n] = gsx * q.s.deta[n];
    c.ss[2 * n + 1] = gsy * q.s.deta[n];
    c.rs[2 * n] = 0.5 * (grx * q.s.deta[n] + gsx * q.s.dxi[n]);
    c.rs[2 * n + 1] = 0.5 * (gry * q.s.deta[n] + gsy * q.s.dxi[n]);
  }
  return true;
}

struct MitcM {
  CovRows top, bottom, right, left, center;
  double a_top, a_bottom, a_right, a_left, a_center;
};

int mitc_membrane_data(const double* xl, const double* yl, MitcM& d) {
  if (!cov_rows(xl, yl, 0.0, 1.0, d.top) || !cov_rows(xl, yl, 0.0, -1.0, d.bottom) ||
      !cov_rows(xl, yl, 1.0, 0.0, d.right) || !cov_rows(xl, yl, -1.0, 0.0, d.left) ||
      !cov_rows(xl, yl, 0.0, 0.0, d.center))
    return 1;
  static const int RS[4] = {-1, 1, 1, -1}, SS[4] = {-1, -1, 1, 1};
  double xr[2] = {0, 0}, xs[2] = {0, 0}, xd[2] = {0, 0};
  for (int n = 0; n < 4; ++n) {
    const double p[2] = {xl[n], yl[n]};
    for (int k = 0; k < 2; ++k) {
      xr[k] += 0.25 * RS[n] * p[k];
      xs[k] += 0.25 * SS[n] * p[k];
      xd[k] += 0.25 * (RS[n] * SS[n]) * p[k];
    }
  }
  // matrix [x_r x_s] (columns), its inverse rows give m_r, m_s
  const double det = xr[0] * xs[1] - xs[0] * xr[1];
  const double rel = std::sqrt(xr[0] * xr[0] + xr[1] * xr[1]) * std::sqrt(xs[0] * xs[0] + xs[1] * xs[1]);
  if (rel < 1e-30 || std::fabs(det) < 1e-6 * rel) return 2;
  const double mr[2] = {xs[1] / det, -xs[0] / det};
  const double ms[2] = {-xr[1] / det, xr[0] / det};
  const double cr = xd[0] * mr[0] + xd[1] * mr[1];
  const double cs = xd[0] * ms[0] + xd[1] * ms[1];
  const double dd = cr * cr + cs * cs - 1.0;
  if (std::fabs(dd) < 1e-12) return 3;
  d.a_top = cr * (cr - 1.0) / (2.0 * dd);
  d.a_bottom = cr * (cr + 1.0) / (2.0 * dd);
  d.a_right = cs * (cs - 1.0) / (2.0 * dd);
  d.a_left = cs * (cs + 1.0) / (2.0 * dd);
  d.a_center = 2.0 * cr * cs / dd;
  return 0;
}

// mitc4plus_membrane_B: Bm[3][8]
void mitc_membrane_B(const MitcM& d, const QPoint& q, double Bm[3][8]) {
  const double r = q.xi, s = q.eta;
  double err[8], ess[8], ers[8];
  for (int k = 0; k < 8; ++k) {
    err[k] = 0.5 * (1.0 - 2.0 * d.a_top + s + 2.0 * d.a_top * s * s) * d.top.rr[k] +
             0.5 * (1.0 - 2.0 * d.a_bottom - s + 2.0 * d.a_bottom * s * s) * d.bottom.rr[k] +
             d.a_right * (-1.0 + s * s) * d.right.ss[k] + d.a_left * (-1.0 + s * s) * d.left.ss[k] +
             d.a_center * (-1.0 + s * s) * d.center.rs[k];
    ess[k] = d.a_top * (-1.0 + r * r) * d.top.rr[k] + d.a_bottom * (-1.0 + r * r) * d.bottom.rr[k] +
             0.5 * (1.0 - 2.0 * d.a_right + r + 2.0 * d.a_right * r * r) * d.right.ss[k] +
             0.5 * (1.0 - 2.0 * d.a_left - r + 2.0 * d.a_left * r * r) * d.left.ss[k] +
             d.a_center * (-1.0 + r * r) * d.center.rs[k];
    ers[k] = 0.25 * (r + 4.0 * d.a_top * r * s) * d.top.rr[k] +
             0.25 * (-r + 4.0 * d.a_bottom * r * s) * d.bottom.rr[k] +
             0.25 * (s + 4.0 * d.a_right * r * s) * d.right.ss[k] +
             0.25 * (-s + 4.0 * d.a_left * r * s) * d.left.ss[k] +
             (1.0 + d.a_center * r * s) * d.center.rs[k];
  }
  const double a = q.Jinv[0][0], b = q.Jinv[1][0], c = q.Jinv[0][1], dd = q.Jinv[1][1];
  for (int k = 0; k < 8; ++k) {
    Bm[0][k] = a * a * err[k] + c * c * ess[k] + 2.0 * a * c * ers[k];
    Bm[1][k] = b * b * err[k] + dd * dd * ess[k] + 2.0 * b * dd * ers[k];
    Bm[2][k] = 2.0 * a * b * err[k] + 2.0 * c * dd * ess[k] + 2.0 * (a * dd + b * c) * ers[k];
  }
}

void std_membrane_B(const QPoint& q, double Bm[3][8]) {
  std::memset(Bm, 0, sizeof(double) * 24);
  for (int n = 0; n < 4; ++n) {
    Bm[0][2 * n] = q.dNdx[0][n];
    Bm[1][2 * n + 1] = q.dNdx[1][n];
    Bm[2][2 * n] = q.dNdx[1][n];
    Bm[2][2 * n + 1] = q.dNdx[0][n];
  }
}

void bending_B(const QPoint& q, double Bb[3][12]) {
  std::memset(Bb, 0, sizeof(double) * 36);
  for (int n = 0; n < 4; ++n) {
    Bb[0][3 * n] = q.dNdx[0][n];
    Bb[1][3 * n + 1] = q.dNdx[1][n];
    Bb[2][3 * n] = q.dNdx[1][n];
    Bb[2][3 * n + 1] = q.dNdx[0][n];
  }
}

void mindlin_shear_B(const QPoint& q, double Bs[2][12]) {
  std::memset(Bs, 0, sizeof(double) * 24);
  for (int n = 0; n < 4; ++n) {
    Bs[0][3 * n + 2] = q.dNdx[0][n];
    Bs[0][3 * n] = -q.s.N[n];
    Bs[1][3 * n + 2] = q.dNdx[1][n];
    Bs[1][3 * n + 1] = -q.s.N[n];
  }
}

struct MitcS {
  double xi_bottom[12], xi_top[12], eta_right[12], eta_left[12], xi_center[12], eta_center[12];
};

void cov_shear(const double* xl, const double* yl, double xi0, double eta0, double* Bxi,
               double* Beta) {
  const QShape s = qshape(xi0, eta0);
  double gxx = 0, gxy = 0, gex = 0, gey = 0;
  for (int n = 0; n < 4; ++n) {
    gxx += s.dxi[n] * xl[n];
    gxy += s.dxi[n] * yl[n];
    gex += s.deta[n] * xl[n];
    gey += s.deta[n] * yl[n];
  }
  for (int n = 0; n < 4; ++n) {
    Bxi[3 * n + 2] = s.dxi[n];
    Bxi[3 * n] = -s.N[n] * gxx;
    Bxi[3 * n + 1] = -s.N[n] * gxy;
    Beta[3 * n + 2] = s.deta[n];
    Beta[3 * n] = -s.N[n] * gex;
    Beta[3 * n + 1] = -s.N[n] * gey;
  }
}

void mitc_shear_data(const double* xl, const double* yl, MitcS& d) {
  double tmp[12];
  cov_shear(xl, yl, 0.0, -1.0, d.xi_bottom, tmp);
  cov_shear(xl, yl, 0.0, 1.0, d.xi_top, tmp);
  cov_shear(xl, yl, 1.0, 0.0, tmp, d.eta_right);
  cov_shear(xl, yl, -1.0, 0.0, tmp, d.eta_left);
  cov_shear(xl, yl, 0.0, 0.0, d.xi_center, d.eta_center);
}

void mitc_shear_B(const MitcS& d, const QPoint& q, double Bs[2][12]) {
  double cxi[12], ceta[12];
  for (int k = 0; k < 12; ++k) {
    cxi[k] = 0.5 * (1.0 - q.eta) * d.xi_bottom[k] + 0.5 * (1.0 + q.eta) * d.xi_top[k] +
             q.xi * (d.xi_center[k] - 0.5 * (d.xi_bottom[k] + d.xi_top[k]));
    ceta[k] = 0.5 * (1.0 - q.xi) * d.eta_left[k] + 0.5 * (1.0 + q.xi) * d.eta_right[k] +
              q.eta * (d.eta_center[k] - 0.5 * (d.eta_left[k] + d.eta_right[k]));
  }
  for (int k = 0; k < 12; ++k) {
    Bs[0][k] = q.Jinv[0][0] * cxi[k] + q.Jinv[0][1] * ceta[k];
    Bs[1][k] = q.Jinv[1][0] * cxi[k] + q.Jinv[1][1] * ceta[k];
  }
}

// K[r][c] += sum B^T D B * w for B (m x nc), D (m x m)
template <int M, int NC>
void btdb(const double (&B)[M][NC], const double (&D)[M][M], double w, double (&K)[NC][NC]) {
  double DB[M][NC];
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < NC; ++j) {
      double s = 0.0;
      for (int k = 0; k < M; ++k) s += D[i][k] * B[k][j];
      DB[i][j] = s;
    }
  for (int i = 0; i < NC; ++i)
    for (int j = 0; j < NC; ++j) {
      double s = 0.0;
      for (int k = 0; k < M; ++k) s += B[k][i] * DB[k][j];
      K[i][j] += s * w;
    }
}


}  // namespace

extern "C" void orc_internal_set_error(const char* msg);

namespace {
// deriv 0: K_e; deriv 1: dK_e/dt (membrane and shear scale with t, bending
// with t^3, the drilling penalty with the rotational diagonal it is taken from).
int quad4_ke_impl(const double* X, double E, double nu, double G, double t, double twelveI_t3,
                  double tst, int form, int deriv, double* Ke_out);
}  // namespace

extern "C" {

int orc_quad4_ke(const double* X, double E, double nu, double G, double t, double twelveI_t3,
                 double tst, int form, double* Ke_out) {
  return quad4_ke_impl(X, E, nu, G, t, twelveI_t3, tst, form, 0, Ke_out);
}

int orc_quad4_dke_dt(const double* X, double E, double nu, double G, double t, double twelveI_t3,
                     double tst, int form, double* dKe_out) {
  return quad4_ke_impl(X, E, nu, G, t, twelveI_t3, tst, form, 1, dKe_out);
}

}  // extern "C"

namespace {
int quad4_ke_impl(const double* X, double E, double nu, double G, double t, double twelveI_t3,
                  double tst, int form, int deriv, double* Ke_out) {
  // frame
  const V3 g[4] = {{X[0], X[1], X[2]}, {X[3], X[4], X[5]}, {X[6], X[7], X[8]}, {X[9], X[10], X[11]}};
  const V3 v12 = sub(g[1], g[0]), v14 = sub(g[3], g[0]);
  const V3 e3 = normalized(cross(v12, v14));
  const V3 e1 = normalized(v12);
  const V3 e2 = cross(e3, e1);
  double xl[4], yl[4];
  for (int n = 0; n < 4; ++n) {
    xl[n] = dot(g[n], e1);
    yl[n] = dot(g[n], e2);
  }
  const double R[3][3] = {{e1.x, e1.y, e1.z}, {e2.x, e2.y, e2.z}, {e3.x, e3.y, e3.z}};
  const double Rr[3][3] = {{-R[1][0], -R[1][1], -R[1][2]}, {R[0][0], R[0][1], R[0][2]},
                           {R[2][0], R[2][1], R[2][2]}};
  // section
  const double c = E / (1.0 - nu * nu);
  const double Dp[3][3] = {{c, c * nu, 0.0}, {c * nu, c, 0.0}, {0.0, 0.0, c * (1.0 - nu) / 2.0}};
  const double Gs = (G > 0.0) ? G : E / (2.0 * (1.0 + nu));
  const double ib = twelveI_t3 * t * t * t / 12.0;
  double A[3][3], Db[3][3], Ts[2][2] = {{tst * t * Gs, 0.0}, {0.0, tst * t * Gs}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      A[i][j] = t * Dp[i][j];
      Db[i][j] = ib * Dp[i][j];
    }
  double An[3][3], Ash[3][3] = {};
  std::memcpy(An, A, sizeof A);
  An[2][2] = 0.0;
  Ash[2][2] = A[2][2];
  MitcM mm{};
  MitcS ms{};
  if (form == 0) {
    const int rc = mitc_membrane_data(xl, yl, mm);
    if (rc == 1) { orc_internal_set_error("CQUAD4: negative Jacobian"); return ORC_ERR_SOLVER; }
    if (rc == 2) { orc_internal_set_error("CQUAD4: MITC4+ characteristic geometry is singular"); return ORC_ERR_SOLVER; }
    if (rc == 3) { orc_internal_set_error("CQUAD4: MITC4+ distortion denominator is near zero"); return ORC_ERR_SOLVER; }
    mitc_shear_data(xl, yl, ms);
  }
  double Kl[24][24] = {};
  double dKl[24][24] = {};  // d/dt of the membrane, bending and shear parts
  static const int MEM[2] = {0, 1}, PLT[3] = {3, 4, 2};
  const double gp = 1.0 / std::sqrt(3.0), gg[2] = {-gp, gp};
  for (int gi = 0; gi < 2; ++gi)
    for (int gj = 0; gj < 2; ++gj) {
      QPoint q;
      if (!eval_point(xl, yl, gg[gi], gg[gj], q)) {
        orc_internal_set_error("CQUAD4: negative Jacobian");
        return ORC_ERR_SOLVER;
      }
      const double w = q.detJ;
      double Bm[3][8], Bb[3][12], Bs[2][12];
      if (form == 0) mitc_membrane_B(mm, q, Bm); else std_membrane_B(q, Bm);
      bending_B(q, Bb);
      if (form == 0) mitc_shear_B(ms, q, Bs); else mindlin_shear_B(q, Bs);
      double Kmm[8][8] = {}, Kbb[12][12] = {}, Kss[12][12] = {};
      btdb<3, 8>(Bm, form == 0 ? A : An, w, Kmm);
      btdb<3, 12>(Bb, Db, w, Kbb);
      btdb<2, 12>(Bs, Ts, w, Kss);
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
          for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) Kl[6 * i + MEM[a]][6 * j + MEM[b]] += Kmm[2 * i + a][2 * j + b];
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) Kl[6 * i + PLT[a]][6 * j + PLT[b]] += Kbb[3 * i + a][3 * j + b];
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) Kl[6 * i + PLT[a]][6 * j + PLT[b]] += Kss[3 * i + a][3 * j + b];
          for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
              dKl[6 * i + MEM[a]][6 * j + MEM[b]] += Kmm[2 * i + a][2 * j + b] / t;
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
              dKl[6 * i + PLT[a]][6 * j + PLT[b]] +=
                  (3.0 * Kbb[3 * i + a][3 * j + b] + Kss[3 * i + a][3 * j + b]) / t;
        }
    }
  if (form == 1 && std::fabs(A[2][2]) > 0.0) {
    QPoint q0;
    eval_point(xl, yl, 0.0, 0.0, q0);
    double Bm0[3][8], Kxy[8][8] = {};
    std_membrane_B(q0, Bm0);
    btdb<3, 8>(Bm0, Ash, q0.detJ * 4.0, Kxy);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j)
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b) {
            Kl[6 * i + MEM[a]][6 * j + MEM[b]] += Kxy[2 * i + a][2 * j + b];
            dKl[6 * i + MEM[a]][6 * j + MEM[b]] += Kxy[2 * i + a][2 * j + b] / t;
          }
  }
  double maxd = 0.0;
  for (int i = 0; i < 24; ++i) maxd = std::max(maxd, std::fabs(Kl[i][i]));
  if (maxd < 1e-10) {
    orc_internal_set_error("CQUAD4: stiffness matrix diagonal is near-zero");
    return ORC_ERR_SOLVER;
  }
  double maxrot = 0.0;
  int krot = -1, kmax = 0;
  for (int i = 0; i < 24; ++i)
    if (std::fabs(Kl[i][i]) > std::fabs(Kl[kmax][kmax])) kmax = i;
  for (int i = 0; i < 4; ++i)
    for (int c = 3; c <= 4; ++c)
      if (std::fabs(Kl[6 * i + c][6 * i + c]) > maxrot) {
        maxrot = std::fabs(Kl[6 * i + c][6 * i + c]);
        krot = 6 * i + c;
      }
  const double drill = (maxrot > 0.0) ? 1e-6 * maxrot : 1e-12 * maxd;
  for (int i = 0; i < 4; ++i) Kl[6 * i + 5][6 * i + 5] += drill;
  if (deriv) {
    const int k = (maxrot > 0.0) ? krot : kmax;
    const double sg = (Kl[k][k] < 0.0) ? -1.0 : 1.0;
    const double ddrill = ((maxrot > 0.0) ? 1e-6 : 1e-12) * sg * dKl[k][k];
    for (int i = 0; i < 4; ++i) dKl[6 * i + 5][6 * i + 5] += ddrill;
    std::memcpy(Kl, dKl, sizeof Kl);
  }
  // T^T Kl T with T = blockdiag(R, Rr) per node
  double T[24][24] = {};
  for (int n = 0; n < 4; ++n)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        T[6 * n + a][6 * n + b] = R[a][b];
        T[6 * n + 3 + a][6 * n + 3 + b] = Rr[a][b];
      }
  double KT[24][24];
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) {
      double s = 0.0;
      for (int k = 0; k < 24; ++k) s += Kl[i][k] * T[k][j];
      KT[i][j] = s;
    }
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) {
      double s = 0.0;
      for (int k = 0; k < 24; ++k) s += T[k][i] * KT[k][j];
      Ke_out[i * 24 + j] = s;
    }
  return ORC_OK;
}
}  // namespace

extern "C" {

int orc_quad4_me(const double* X, double rho, double t, double* Me_out) {
  const V3 g[4] = {{X[0], X[1], X[2]}, {X[3], X[4], X[5]}, {X[6], X[7], X[8]}, {X[9], X[10], X[11]}};
  const V3 v12 = sub(g[1], g[0]), v14 = sub(g[3], g[0]);
  const V3 e3 = normalized(cross(v12, v14));
  const V3 e1 = normalized(v12);
  const V3 e2 = cross(e3, e1);
  double xl[4], yl[4];
  for (int n = 0; n < 4; ++n) {
    xl[n] = dot(g[n], e1);
    yl[n] = dot(g[n], e2);
  }
  const double mpa = rho * t, rpa = rho * t * t * t / 12.0;
  std::fill(Me_out, Me_out + 576, 0.0);
  if (mpa == 0.0 && rpa == 0.0) return ORC_OK;
  const double R[3][3] = {{e1.x, e1.y, e1.z}, {e2.x, e2.y, e2.z}, {e3.x, e3.y, e3.z}};
  const double Rr[3][3] = {{-R[1][0], -R[1][1], -R[1][2]}, {R[0][0], R[0][1], R[0][2]},
                           {R[2][0], R[2][1], R[2][2]}};
  const double drill_scale = 1e-4 * rpa / 100.0;
  double Ml[24][24] = {};
  const double gp = 1.0 / std::sqrt(3.0), gg[2] = {-gp, gp};
  for (int gi = 0; gi < 2; ++gi)
    for (int gj = 0; gj < 2; ++gj) {
      QPoint q;
      if (!eval_point(xl, yl, gg[gi], gg[gj], q)) return ORC_ERR_SOLVER;
      for (int a = 0; a < 4; ++a) {
        for (int b = 0; b < 4; ++b) {
          const double mab = q.s.N[a] * q.s.N[b] * q.detJ;
          for (int d = 0; d < 3; ++d) Ml[6 * a + d][6 * b + d] += mpa * mab;
          for (int d = 0; d < 2; ++d) Ml[6 * a + 3 + d][6 * b + 3 + d] += rpa * mab;
        }
        Ml[6 * a + 5][6 * a + 5] += drill_scale * q.s.N[a] * q.s.N[a] * q.detJ;
      }
    }
  double T[24][24] = {};
  for (int n = 0; n < 4; ++n)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        T[6 * n + a][6 * n + b] = R[a][b];
        T[6 * n + 3 + a][6 * n + 3 + b] = Rr[a][b];
      }
  double MT[24][24];
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) {
      double s = 0.0;
      for (int k = 0; k < 24; ++k) s += Ml[i][k] * T[k][j];
      MT[i][j] = s;
    }
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) {
      double s = 0.0;
      for (int k = 0; k < 24; ++k) s += T[k][i] * MT[k][j];
      Me_out[i * 24 + j] = s;
    }
  return ORC_OK;
}

// recover_shell_response (cquad4.cpp:758-801) at (xi, eta), no thermal strain:
// u_local = T u (to_local_displacements 491-502), membrane DOFs (u, v) and
// plate DOFs (rx, ry, w) of each node (gather_* 504-522), membrane strain
// Bm u_m (MITC4+ tied B, or the standard B for Mindlin), curvature Bb u_p.
// out6 = membrane stress Dp eps (sx, sy, sxy) then moments D kappa (mx, my,
// mxy), D = (12I/t^3) t^3/12 Dp (build_shell_section 189-218; no MID4
// coupling, so the section's B term is zero).
int orc_quad4_recover(const double* X, double E, double nu, double t, double twelveI_t3,
                      int form, const double* ue, double xi, double eta, double* out6) {
  const V3 g[4] = {{X[0], X[1], X[2]}, {X[3], X[4], X[5]}, {X[6], X[7], X[8]}, {X[9], X[10], X[11]}};
  const V3 v12 = sub(g[1], g[0]), v14 = sub(g[3], g[0]);
  const V3 e3 = normalized(cross(v12, v14));
  const V3 e1 = normalized(v12);
  const V3 e2 = cross(e3, e1);
  double xl[4], yl[4];
  for (int n = 0; n < 4; ++n) {
    xl[n] = dot(g[n], e1);
    yl[n] = dot(g[n], e2);
  }
  const double R[3][3] = {{e1.x, e1.y, e1.z}, {e2.x, e2.y, e2.z}, {e3.x, e3.y, e3.z}};
  const double Rr[3][3] = {{-R[1][0], -R[1][1], -R[1][2]}, {R[0][0], R[0][1], R[0][2]},
                           {R[2][0], R[2][1], R[2][2]}};
  double ul[24];
  for (int n = 0; n < 4; ++n)
    for (int a = 0; a < 3; ++a) {
      double st = 0.0, sr = 0.0;
      for (int b = 0; b < 3; ++b) {
        st += R[a][b] * ue[6 * n + b];
        sr += Rr[a][b] * ue[6 * n + 3 + b];
      }
      ul[6 * n + a] = st;
      ul[6 * n + 3 + a] = sr;
    }
  double um[8], up[12];
  for (int n = 0; n < 4; ++n) {
    um[2 * n] = ul[6 * n];
    um[2 * n + 1] = ul[6 * n + 1];
    up[3 * n] = ul[6 * n + 3];
    up[3 * n + 1] = ul[6 * n + 4];
    up[3 * n + 2] = ul[6 * n + 2];
  }
  QPoint q;
  if (!eval_point(xl, yl, xi, eta, q)) {
    orc_internal_set_error("CQUAD4: negative Jacobian");
    return ORC_ERR_SOLVER;
  }
  double Bm[3][8], Bb[3][12];
  if (form == 0) {
    MitcM mm{};
    const int rc = mitc_membrane_data(xl, yl, mm);
    if (rc == 1) { orc_internal_set_error("CQUAD4: negative Jacobian"); return ORC_ERR_SOLVER; }
    if (rc == 2) { orc_internal_set_error("CQUAD4: MITC4+ characteristic geometry is singular"); return ORC_ERR_SOLVER; }
    if (rc == 3) { orc_internal_set_error("CQUAD4: MITC4+ distortion denominator is near zero"); return ORC_ERR_SOLVER; }
    mitc_membrane_B(mm, q, Bm);
  } else {
    std_membrane_B(q, Bm);
  }
  bending_B(q, Bb);
  double eps[3], kap[3];
  for (int i = 0; i < 3; ++i) {
    double s = 0.0, k = 0.0;
    for (int j = 0; j < 8; ++j) s += Bm[i][j] * um[j];
    for (int j = 0; j < 12; ++j) k += Bb[i][j] * up[j];
    eps[i] = s;
    kap[i] = k;
  }
  const double c = E / (1.0 - nu * nu);
  const double Dp[3][3] = {{c, c * nu, 0.0}, {c * nu, c, 0.0}, {0.0, 0.0, c * (1.0 - nu) / 2.0}};
  const double ib = twelveI_t3 * t * t * t / 12.0;
  for (int i = 0; i < 3; ++i) {
    double s = 0.0, m = 0.0;
    for (int j = 0; j < 3; ++j) {
      s += Dp[i][j] * eps[j];
      m += ib * Dp[i][j] * kap[j];
    }
    out6[i] = s;
    out6[3 + i] = m;
  }
  return ORC_OK;
}

}  // extern "C"
