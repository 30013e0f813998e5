// vibestran_amd/csrc/vbh_elem.hpp -- pieces shared by the element kernels
// (solids in vbh_assemble.hip, CQUAD4 in vbh_shell.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace vbh {

// Lower-triangle (i >= j) pair list of an N x N element matrix, row-major:
// entry t packs (i << 8) | j. The first M(M+1)/2 entries of any table with
// N >= M enumerate the M x M triangle.
template <int N>
struct PairTable {
  unsigned short v[N * (N + 1) / 2];
  constexpr PairTable() : v{} {
    int t = 0;
    for (int i = 0; i < N; ++i)
      for (int j = 0; j <= i; ++j) v[t++] = static_cast<unsigned short>((i << 8) | j);
  }
};

constexpr int kGdStride = 24;   // element DOF list stride (max DOFs per element)
constexpr int kElemWaves = 4;   // waves per block in the element kernels

// Error word written by the element kernels: err[0] bit set, err[1] lowest
// offending element index (atomicMin).
enum : int {
  kErrJacobian = 1,      // non-positive Jacobian / volume
  kErrKaa = 2,           // EAS enhanced block not SPD
  kErrSlot = 4,          // scatter target missing from the pattern (internal)
  kErrMitcGeometry = 32, // CQUAD4 MITC4+ characteristic geometry singular
  kErrMitcDistort = 64,  // CQUAD4 MITC4+ distortion denominator near zero
  kErrShellDiag = 128,   // CQUAD4 stiffness diagonal near zero
};

// Binary search of column `col` in CSR row `row`; -1 if absent.
__device__ __forceinline__ int find_slot(const int* __restrict__ rp, const int* __restrict__ ci,
                                         int row, int col) {
  int lo = rp[row], hi = rp[row + 1] - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int c = ci[mid];
    if (c == col) return mid;
    if (c < col) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

// Batched CQUAD4 stiffness: element s's 300 lower entries (pair order of
// PairTable<24>) at ke[s * 300 ..]; `form` 0 = MITC4+, 1 = Mindlin.
void launch_quad4_stiffness(int nsel, const int* sel, const int* eptr, const int* conn,
                            const double* xyz, const double* E, const double* nu,
                            const double* thick, int form, const int* gd, double* ke, int* err,
                            hipStream_t s);

// out[e] = -u_e^T (dK_e/dp_e) v_e for the selected CQUAD4 elements; param 0:
// p = E, 1: p = thickness.
void launch_quad4_sensitivity(int nsel, const int* sel, const int* eptr, const int* conn,
                              const double* xyz, const double* E, const double* nu,
                              const double* thick, int form, const int* gd, int param,
                              const double* u, const double* v, double* out, int* err,
                              hipStream_t s);

// Internal forces K_e u_e of the selected CQUAD4 elements added to
// fout[6 * node + dof] (basic frame).
void launch_quad4_forces(int nsel, const int* sel, const int* eptr, const int* conn,
                         const double* xyz, const double* E, const double* nu, const double* thick,
                         int form, const int* gd, const double* u, double* fout, int* err,
                         hipStream_t s);

// Stress recovery of the selected CQUAD4 elements: centroid rows cen[e*10..]
// and, when cor is not null, corner rows cor[e*80 + 10*c..] (c < 4; rows 4..7
// are zeroed). Columns [sx, sy, sxy, sz, syz, szx, mx, my, mxy, von_mises].
void launch_quad4_recover(int nsel, const int* sel, const int* eptr, const int* conn,
                          const double* xyz, const double* E, const double* nu,
                          const double* thick, int form, const int* gd, const double* u,
                          double* cen, double* cor, int* err, hipStream_t s);

}  // namespace vbh
