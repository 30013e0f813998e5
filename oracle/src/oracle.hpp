// oracle/src/oracle.hpp
//
// TEST INFRASTRUCTURE ONLY. This is the CPU FP64 restatement of the
// reference's linear-static hot path (vibestran). It is the checker for the
// HIP product path: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it. The product library never links or calls it.
//
// Every function cites the reference file:line it restates. The reference
// itself cannot be compiled here (Eigen, spdlog, CHOLMOD and <format> are
// absent), so these restatements are pinned by the reference's own unit-test
// known answers, e2e decks and closed forms instead (see tests/).
#pragma once
#include <cstdint>

extern "C" {

// Status codes returned by every orc_* entry point.
enum { ORC_OK = 0, ORC_ERR_ARG = 1, ORC_ERR_SOLVER = 3 };

// Thread-local text of the last error (reference SolverError::what()).
const char* orc_last_error(void);

// ---- DofMap: src/core/dof_map.cpp:10-104 --------------------------------
// Opaque map over the node set; slots are 6 per node in sorted-ID order.
typedef struct OrcDofMap OrcDofMap;
OrcDofMap* orc_dofmap_build(int nnodes, const int* node_ids, int dofs_per_node);
void orc_dofmap_free(OrcDofMap* m);
int  orc_dofmap_constrain(OrcDofMap* m, int node, int dof0);
int  orc_dofmap_constrain_batch(OrcDofMap* m, int count, const int* nodes, const int* dofs);
int  orc_dofmap_eq_index(const OrcDofMap* m, int node, int dof0);
int  orc_dofmap_num_free(const OrcDofMap* m);
int  orc_dofmap_num_total(const OrcDofMap* m);

// Element type codes shared by the oracle's model-level helpers.
enum { ORC_CHEXA8 = 1, ORC_CPENTA6 = 2, ORC_CTETRA4 = 3, ORC_CQUAD4 = 4 };

// build_analysis_dof_map (src/solver/analysis_support.cpp:176-214): activate
// DOFs per element family (39-68), constrain inactive then SPC DOFs.
// spc_masks[i] bit d (0..5) set = DOF d+1 of spc_nodes[i] is constrained.
// eq_out is nnodes*6, in the order of node_ids. Returns n_free or <0.
int orc_analysis_dof_map(int nnodes, const int* node_ids,
                         int nelem, const int* elem_type, const int* elem_ptr,
                         const int* elem_nodes,
                         int nspc, const int* spc_nodes, const int* spc_masks,
                         int* eq_out);

// ---- Elements ------------------------------------------------------------
// All element matrices are written row-major, ndof x ndof.
// CHexa8 SRI: src/elements/solid_elements.cpp:95-180
int orc_hexa8_ke_sri(const double* xyz24, double E, double nu, double* ke576);
// CHexa8Eas: src/elements/solid_elements.cpp:1054-1206
int orc_hexa8_ke_eas(const double* xyz24, double E, double nu, double* ke576);
// CTetra4: src/elements/solid_elements.cpp:312-384
int orc_tetra4_ke(const double* xyz12, double E, double nu, double* ke144);
// CPenta6Eas: src/elements/solid_elements.cpp:1333-1467
int orc_penta6_ke_eas(const double* xyz18, double E, double nu, double* ke324);
// Consistent mass (CHexa8 182-217 / CHexa8Eas 1208-1243), CTetra4 386-411.
int orc_hexa8_me(const double* xyz24, double rho, double* me576);
int orc_tetra4_me(const double* xyz12, double rho, double* me144);
int orc_penta6_me(const double* xyz18, double rho, double* me324);

// ---- SparseMatrixBuilder / CsrData: src/core/sparse_matrix.cpp -----------
typedef struct OrcBuilder OrcBuilder;
typedef struct OrcCsr OrcCsr;
OrcBuilder* orc_builder_create(int n);
void orc_builder_free(OrcBuilder* b);
void orc_builder_add(OrcBuilder* b, int row, int col, double v);          // hpp:43-45
int  orc_builder_add_element(OrcBuilder* b, int ndof, const int* gdofs,
                             const double* ke_rowmajor);                  // cpp:16-35
long long orc_builder_num_triplets(const OrcBuilder* b);
OrcCsr* orc_builder_build_csr(OrcBuilder* b);                              // cpp:62-150
void orc_csr_free(OrcCsr* c);
int  orc_csr_n(const OrcCsr* c);
int  orc_csr_nnz(const OrcCsr* c);
int  orc_csr_is_lower(const OrcCsr* c);
void orc_csr_copy(const OrcCsr* c, int* row_ptr, int* col_ind, double* values);
OrcCsr* orc_csr_from_arrays(int n, int nnz, const int* row_ptr, const int* col_ind,
                            const double* values, int lower_only);
int  orc_csr_multiply(const OrcCsr* c, const double* x, double* y);          // cpp:152-181
OrcCsr* orc_csr_expanded_symmetric(const OrcCsr* c);                         // cpp:183-235
OrcCsr* orc_csr_lower_triangle(const OrcCsr* c);                             // cpp:237-274

// Model-level assembly of a solid mesh (assembly_parallel.hpp:68-110 serial
// branch, element_factory.cpp:31-61). conn uses the element's node count
// (8 CHEXA8 / 6 CPENTA6 / 4 CTETRA4) with 0-based node indices into xyz;
// eq is the nnodes*6 analysis DOF map. formulation: 0 = EAS (default),
// 1 = SRI (PSOLID ISOP=SRI; CTETRA4 ignores it).
OrcCsr* orc_assemble_solids(int n, int nelem, const int* elem_type, const int* elem_ptr,
                            const int* conn, const double* xyz, const int* eq,
                            const double* E, const double* nu, int formulation);

// ---- Jacobi-PCG (semantics of cuda_pcg_backend_impl.hpp:1689-1999 Jacobi
// branch with build_jacobi_diagonal 2076-2101 and validate 2103-2114).
// Returns ORC_OK, or ORC_ERR_SOLVER with orc_last_error() text.
int orc_pcg_jacobi(const OrcCsr* K, const double* F, double tol, int max_iters,
                   int nthreads, double* u, int* iters, double* rel_res);
// Same recurrence, timed for a fixed number of iterations (CPU baseline).
int orc_pcg_jacobi_fixed(const OrcCsr* K, const double* F, int iters, int nthreads,
                         double* u, double* seconds);
// True relative residual ||F-Ku||/||F|| (cuda_pcg_backend_impl.hpp:90-114).
double orc_true_relative_residual(const OrcCsr* K, const double* u, const double* F);

// ---- Loads -----------------------------------------------------------------
// PLOAD4 on a CHEXA8 face: face selection select_solid_face
// (linear_static_pressure.cpp:430-517) and 2x2 Gauss integration
// integrate_quad_surface_load (154-196) with normal_sign -1 (solids, 745-753).
// g1_local/g34_local are element-local corner indices (or -1 for none);
// pressures are the 4 corner pressures. forces24 receives the element
// force vector [3 per element node].
int orc_pload4_hexa8(const double* xyz24, int g1_local, int g34_local,
                     const double* pressures4, double* forces24);

}  // extern "C"

extern "C" {
// ---- CQUAD4 shell: compute_shell_stiffness (src/elements/cquad4.cpp:580-670)
// with compute_shell_frame 131-156, build_shell_section 189-218 (MID2/MID3 =
// MID1, no MID4 coupling), evaluate_quad_point 220-252, MITC4+ membrane
// 311-415 and shear 417-489, Mindlin membrane/shear 254-285 + one-point
// in-plane shear 647-655, drilling penalty 109-129. formulation: 0 MITC4+
// (default), 1 Mindlin. G <= 0 selects E/(2(1+nu)). Output row-major 24x24.
int orc_quad4_ke(const double* xyz12, double E, double nu, double G, double t,
                 double twelveI_t3, double tst, int formulation, double* ke576);
// Analytic dK_e/dt of orc_quad4_ke (adjoint sensitivities; no reference
// counterpart): membrane and transverse shear are linear in t, bending cubic,
// the drilling penalty follows the rotational diagonal entry it is taken from.
int orc_quad4_dke_dt(const double* xyz12, double E, double nu, double G, double t,
                     double twelveI_t3, double tst, int formulation, double* dke576);
// compute_shell_mass (cquad4.cpp:672-708): rho*t translational, rho*t^3/12
// rotary, tiny drilling inertia.
int orc_quad4_me(const double* xyz12, double rho, double t, double* me576);
// Mixed solid/shell assembly: like orc_assemble_solids plus CQUAD4 elements
// (6 DOFs per node, thickness[e]).
OrcCsr* orc_assemble_model(int n, int nelem, const int* elem_type, const int* elem_ptr,
                           const int* conn, const double* xyz, const int* eq, const double* E,
                           const double* nu, const double* thickness, int solid_form,
                           int shell_form);
// GRAV load vector (LinearStaticSolver::apply_inertial_loads,
// linear_static.cpp:400-503): F = sum_e M_e a with the consistent element
// masses, acceleration g[3] at every node (basic frame, WTMASS folded into rho).
int orc_grav_rhs(int n, int nelem, const int* elem_type, const int* elem_ptr, const int* conn,
                 const double* xyz, const int* eq, const double* rho, const double* thickness,
                 const double* g, double* F);
// Internal nodal forces f[6 node + dof] = sum_e K_e u_e (no reference
// counterpart): every DOF of every node, constrained ones included.
int orc_element_forces(int nelem, const int* elem_type, const int* elem_ptr, const int* conn,
                       const double* xyz, const int* eq, const double* E, const double* nu,
                       const double* thickness, int solid_form, int shell_form, const double* u,
                       int nnodes, double* f);
// recover_shell_response (cquad4.cpp:758-801) at (xi, eta) without thermal
// strain: out6 = membrane stress (sx, sy, sxy) and moments (mx, my, mxy) from
// the 24 global element displacements ue.
int orc_quad4_recover(const double* xyz12, double E, double nu, double t, double twelveI_t3,
                      int formulation, const double* ue24, double xi, double eta, double* out6);
// Element stress recovery (LinearStaticSolver::recover_results element part,
// linear_static.cpp:676-1197): centroid row per element and up to 8 corner
// rows (element node order; unused rows 0), each
// [sx, sy, sxy, sz, syz, szx, mx, my, mxy, von_mises] as in the element CSV
// (csv_writer.cpp:66-85). Solids use the displacement B at the point, also for
// EAS elements. centroid is nelem*10, corner nelem*80 or null.
int orc_recover_stress(int nelem, const int* elem_type, const int* elem_ptr, const int* conn,
                       const double* xyz, const int* eq, const double* E, const double* nu,
                       const double* thickness, int shell_form, const double* u,
                       double* centroid, double* corner);
// FP64 CPU adjoint contraction (no reference counterpart): out[e] =
// -u_e^T (dK_e/dp_e) v_e, p = E (param 0) or CQUAD4 thickness (param 1;
// solids 0). Compliance: v = u; a functional L^T u: v = lambda, K lambda = L.
int orc_element_sensitivity(int n, int nelem, const int* elem_type, const int* elem_ptr,
                            const int* conn, const double* xyz, const int* eq, const double* E,
                            const double* nu, const double* thickness, int solid_form,
                            int shell_form, int param, const double* u, const double* v,
                            double* out);
}
