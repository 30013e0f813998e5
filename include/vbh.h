/* include/vbh.h -- C ABI of the MI355X-native hip-pcg backend (libvbh.so).
 *
 * The library replaces, behind a plain C boundary (no C++ or torch types),
 * the reference's solver plugin and its CSR stiffness assembly:
 *   - vibestran::SolverBackend / CudaPCGSolverBackend
 *       (reference include/solver/solver_backend.hpp:20-51,
 *        include/solver/cuda_pcg_solver_backend.hpp:29-76)
 *   - SparseMatrixBuilder::add_element_stiffness/build_csr driven per element
 *       (reference src/solver/assembly_parallel.hpp:68-110,
 *        src/core/sparse_matrix.cpp:16-150)
 *   - ElementBase::stiffness_matrix for CHEXA8 (EAS default, SRI), CPENTA6
 *     (EAS), CTETRA4 and CQUAD4 (MITC4+ default, Mindlin)
 *       (reference include/elements/element_base.hpp:30-61)
 *
 * Conventions: all host pointers are caller-owned and only read/written for
 * the duration of the call. Every entry point returns a vbh_status; on
 * failure vbh_last_error() holds a thread-local message whose text follows
 * the reference's SolverError messages. Indices are 0-based int32
 * (reference EqIndex, include/core/dof_map.hpp:19). One context is used by
 * one host thread at a time (the reference backend is not thread-safe
 * either, solver_backend.hpp:20-51).
 */
#ifndef VBH_H
#define VBH_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vbh_ctx vbh_ctx;

typedef enum {
  VBH_OK = 0,
  VBH_ERR_ARG = 1,        /* invalid argument (std::invalid_argument in the reference) */
  VBH_ERR_NO_DEVICE = 2,  /* try_create() -> nullopt (cuda_pcg_solver_backend.hpp:45-47) */
  VBH_ERR_SOLVER = 3,     /* SolverError (include/core/exceptions.hpp:14-17) */
  VBH_ERR_HIP = 4,        /* HIP runtime failure */
  VBH_ERR_STATE = 5       /* call order error (e.g. solve before a matrix is set) */
} vbh_status;

/* Precision modes for vbh_solve / vbh_solve_block. */
typedef enum {
  VBH_PREC_FP64 = 0,   /* plain FP64 Jacobi-PCG: cuda-pcg (cuda_pcg_backend_impl.hpp:2116-2132) */
  VBH_PREC_MIXED = 1   /* FP32 SpMV + Jacobi, FP64 true residual and solution update
                          (replaces cuda-pcg-mixed, cuda_pcg_backend_impl.hpp:2134-2262) */
} vbh_precision;

/* Element type codes for the batched assembly entry points. */
typedef enum { VBH_CHEXA8 = 1, VBH_CPENTA6 = 2, VBH_CTETRA4 = 3, VBH_CQUAD4 = 4 } vbh_elem_type;

/* Solid formulations (PSOLID ISOP, reference element_factory.cpp:31-61). */
typedef enum { VBH_SOLID_EAS = 0, VBH_SOLID_SRI = 1 } vbh_solid_form;
/* Shell formulations (PARAM,SHELLFORM, reference element_factory.cpp:13-24). */
typedef enum { VBH_SHELL_MITC4 = 0, VBH_SHELL_MINDLIN = 1 } vbh_shell_form;

/* Diagnostics of the most recent solve (reference last_iteration_count /
 * last_estimated_error, solver_backend.hpp:36-42) plus device timings. */
typedef struct {
  int iterations;         /* PCG iterations (SpMV count of the inner solver) */
  int refinements;        /* FP64 true-residual recomputes (mixed mode) */
  double rel_res;         /* final relative residual ||F-Ku||/||F|| */
  double solve_ms;        /* device time of the solve, hipEvents on the ctx stream */
  double spmv_ms_avg;     /* mean duration of one inner SpMV launch (0 if unmeasured) */
  int format;             /* device operator format: 0 scalar CSR, 3 BSR3, 6 BSR6 */
} vbh_solve_stats;

/* Counters of the most recent vbh_assemble call. */
typedef struct {
  long long nnz_topological;  /* lower pairs implied by element connectivity */
  long long nnz;              /* stored lower nonzeros after exact-zero drops */
  double pattern_ms;          /* device time: pair keys, radix sort, unique */
  double element_ms;          /* device time: element matrices + scatter */
  double mpc_ms;              /* wall time of K_red = T^T K T (0 without MPCs) */
} vbh_assembly_stats;

/* ---- errors / devices ---------------------------------------------------- */
const char* vbh_last_error(void);
int vbh_device_count(void);

/* ---- context: CudaPCGSolverBackend::try_create (cuda_pcg_backend_impl.hpp:2041-2065).
 * tol <= 0 selects 1e-8, max_iters <= 0 selects 10000 (impl:2055).
 * Returns VBH_ERR_NO_DEVICE when no GPU is usable. */
int vbh_ctx_create(int device, double tol, int max_iters, vbh_ctx** out);
void vbh_ctx_destroy(vbh_ctx* ctx);
/* CudaPCGSolverBackend::device_name (cuda_pcg_solver_backend.hpp:72) */
const char* vbh_device_name(const vbh_ctx* ctx);
int vbh_set_precision(vbh_ctx* ctx, int precision);
int vbh_set_tolerance(vbh_ctx* ctx, double tol, int max_iters);
/* Operator format for the SpMV: -1 auto, 0 scalar CSR, 3 BSR 3x3, 6 BSR 6x6. */
int vbh_set_format(vbh_ctx* ctx, int format);

/* ---- matrix input at the SolverBackend boundary ------------------------
 * K as SparseMatrixBuilder::CsrData (sparse_matrix.hpp:63-87): row_ptr[n+1],
 * col_ind[nnz], values[nnz], rows sorted by column, duplicates summed;
 * lower_only != 0 means symmetric lower storage (L + L^T - diag L).
 * Validates like validate_pcg_inputs / build_jacobi_diagonal
 * (cuda_pcg_backend_impl.hpp:2076-2114): empty matrix, zero or missing
 * diagonal -> VBH_ERR_SOLVER. */
int vbh_set_csr(vbh_ctx* ctx, int n, int nnz, const int* row_ptr, const int* col_ind,
                const double* values, int lower_only);
/* Size of the lower CSR currently held (after vbh_set_csr or assembly). */
int vbh_get_csr_size(const vbh_ctx* ctx, int* n, int* nnz);
/* Download the lower CSR held by the context (any pointer may be NULL). */
int vbh_get_csr(const vbh_ctx* ctx, int* row_ptr, int* col_ind, double* values);

/* ---- DOF numbering and assembly on the device ----------------------------
 * vbh_dof_map: build_analysis_dof_map (reference src/solver/analysis_support.cpp:176-214
 * with DofMap::build/constrain_batch, src/core/dof_map.cpp:10-97): 6 slots per
 * node in sorted node-ID order; solids activate T1-T3, CQUAD4 all six
 * (analysis_support.cpp:39-68); SPC'd DOFs removed. conn holds 0-based node
 * indices; spc_masks bit d = DOF d+1. eq_out is nnodes*6 (-1 = constrained). */
int vbh_dof_map(vbh_ctx* ctx, int nnodes, const int* node_ids, int nelem, const int* elem_type,
                const int* elem_ptr, const int* conn, int nspc, const int* spc_nodes,
                const int* spc_masks, int* eq_out, int* n_free);
/* vbh_assemble: element stiffness for every element (CHEXA8 EAS/SRI, CPENTA6
 * EAS, CTETRA4; CQUAD4 MITC4+/Mindlin) and the lower CSR of
 * SparseMatrixBuilder::add_element_stiffness + build_csr (reference
 * src/solver/assembly_parallel.hpp:68-110, src/core/sparse_matrix.cpp:16-150),
 * built and kept on the device as the context's operator. eq is the nnodes*6
 * map of vbh_dof_map; thickness may be NULL for solid-only meshes. */
int vbh_assemble(vbh_ctx* ctx, int n, int nnodes, const double* xyz, const int* eq, int nelem,
                 const int* elem_type, const int* elem_ptr, const int* conn, const double* E,
                 const double* nu, const double* thickness, int solid_form, int shell_form);
int vbh_last_assembly_stats(const vbh_ctx* ctx, vbh_assembly_stats* out);

/* ---- multi-point constraints: MpcHandler (reference src/core/mpc_handler.cpp)
 * vbh_set_mpc: MpcHandler::build (mpc_handler.cpp:14-164) over the numbering
 * eq (nnodes x 6, -1 = constrained, n_full equations, as given to
 * vbh_assemble). Equation k has the terms mpc_ptr[k] .. mpc_ptr[k+1]-1 of
 * term_node (0-based node index), term_dof (1..6) and term_coeff: the term of
 * largest |coefficient| is dependent and eliminated; an equation whose
 * dependent DOF is constrained is skipped (message on stderr); terms on
 * constrained or dependent DOFs drop out; a dependency cycle fails with
 * VBH_ERR_SOLVER. The elimination applies from the next vbh_assemble on:
 * that call assembles K_full (apply_to_stiffness, 188-278) and keeps
 * K_red = T^T K_full T as the operator (n_reduced equations);
 * vbh_assemble_rhs then gives F_red = T^T F (apply_to_force, 280-310), and the
 * recovery, force and sensitivity entry points take u in the reduced numbering
 * and expand it by u_full = T u_red (recover_dependent_dofs, 312-330).
 * nmpc = 0 clears the elimination. */
typedef struct {
  int n_full, n_reduced;
  int n_eliminations;  /* equations eliminated */
  int n_skipped;       /* equations skipped: dependent DOF constrained */
  int t_nnz;           /* entries of T */
} vbh_mpc_stats;
int vbh_set_mpc(vbh_ctx* ctx, int nnodes, const int* eq, int n_full, int nmpc, const int* mpc_ptr,
                const int* term_node, const int* term_dof, const double* term_coeff,
                int* n_reduced);
int vbh_mpc_info(const vbh_ctx* ctx, vbh_mpc_stats* out);
/* T of the elimination set by vbh_set_mpc as CSR over the n_full rows:
 * t_ptr n_full + 1, t_col / t_val t_nnz (any pointer may be NULL). */
int vbh_mpc_transform(const vbh_ctx* ctx, int* t_ptr, int* t_col, double* t_val);
/* u_full (n_full, the mesh numbering) = T u_red over the mesh of the last
 * vbh_assemble (identity copy when that mesh has no elimination). */
int vbh_mpc_expand(vbh_ctx* ctx, const double* u_red, double* u_full);

/* ---- load vector of one load case on the device ----------------------------
 * Over the mesh of the last vbh_assemble (reference LinearStaticSolver load
 * assembly): FORCE/MOMENT (linear_static.cpp:344-398): nodal[6*nnodes]
 * components per node in the basic frame, or NULL; PLOAD4 on solid faces
 * (linear_static_pressure.cpp:154-196): nface faces given as 4 node indices in
 * outward corner order (the face choice of 420-517 is the caller's) and 4
 * corner pressures, or nface = 0; GRAV (linear_static.cpp:400-503): grav[3]
 * acceleration in the basic frame applied to every node with the consistent
 * mass, rho[nelem] densities already scaled by WTMASS, or grav = NULL.
 * F (length n; n_reduced of vbh_set_mpc under an elimination) is overwritten. */
int vbh_assemble_rhs(vbh_ctx* ctx, const double* nodal, int nface, const int* face_nodes,
                     const double* face_p, const double* grav, const double* rho, double* F);
/* Same with F a device pointer (nodal, faces, grav and rho stay host arrays). */
int vbh_assemble_rhs_device(vbh_ctx* ctx, const double* nodal, int nface, const int* face_nodes,
                            const double* face_p, const double* grav, const double* rho,
                            double* dF);

/* ---- adjoint sensitivities (no reference counterpart) ----------------------
 * For every element e of the mesh given to the last vbh_assemble:
 *   out[e] = -u_e^T (dK_e / dp_e) v_e,  p = E_e (param 0) or the CQUAD4
 *   thickness t_e (param 1; solid elements give 0),
 * with u_e, v_e the element's entries of the length-n vectors u, v
 * (constrained DOFs contribute 0). K_e is linear in E; for CQUAD4 membrane and
 * transverse shear scale with t, bending with t^3 and the drilling penalty
 * with the rotational diagonal entry it is taken from. Compliance C = F^T u
 * (K u = F, F independent of p): dC/dp_e = out with v = u. A functional
 * L^T u: solve K lambda = L, then d(L^T u)/dp_e = out with v = lambda.
 * out has nelem entries in element order. */
int vbh_element_sensitivity(vbh_ctx* ctx, int param, const double* u, const double* v,
                            double* out);
/* Same with device-resident u, v and out. */
int vbh_element_sensitivity_device(vbh_ctx* ctx, int param, const double* du, const double* dv,
                                   double* dout);

/* ---- result recovery: LinearStaticSolver::recover_results
 * (linear_static.cpp:566-1201) over the mesh of the last vbh_assemble.
 * u is the solution (n free DOFs). Element rows hold the element CSV columns
 * [sx, sy, sxy, sz, syz, szx, mx, my, mxy, von_mises] (csv_writer.cpp:66-85);
 * components that do not apply to the element are 0. centroid is nelem x 10
 * (solids: centroid stress, 1047-1197; CQUAD4: centroid membrane stress and
 * moments, 712-736). corner (nelem x 8 x 10, may be NULL) holds the per-node
 * points in element node order (CHEXA8 8, CPENTA6 6, CTETRA4 4 copies of the
 * constant stress, CQUAD4 4; unused rows 0). Solids use the displacement B at
 * the point, EAS formulations included, as the reference does. */
int vbh_recover_stress(vbh_ctx* ctx, const double* u, double* centroid, double* corner);
int vbh_recover_stress_device(vbh_ctx* ctx, const double* du, double* dcentroid,
                              double* dcorner);
/* Node displacements T1..R3 (recover_results 568-612, basic frame): disp is
 * nnodes x 6 in the order of the assembled mesh's nodes; constrained and
 * inactive DOFs are 0. */
int vbh_recover_displacements(vbh_ctx* ctx, const double* u, double* disp);
int vbh_recover_displacements_device(vbh_ctx* ctx, const double* du, double* ddisp);
/* Internal nodal forces f = sum_e K_e u_e over the mesh of the last
 * vbh_assemble, at every node DOF (nnodes x 6, basic frame, constrained DOFs
 * included; no reference counterpart). */
int vbh_element_forces(vbh_ctx* ctx, const double* u, double* f);
int vbh_element_forces_device(vbh_ctx* ctx, const double* du, double* df);
/* R = f - F_applied at every node DOF (nnodes x 6) for one load case given as
 * in vbh_assemble_rhs: at constrained DOFs the support reactions
 * (K_full u - F, no reference counterpart), at free DOFs minus the
 * equilibrium residual. Host arrays. */
int vbh_reactions(vbh_ctx* ctx, const double* u, const double* nodal, int nface,
                  const int* face_nodes, const double* face_p, const double* grav,
                  const double* rho, double* R);
/* CSV result files of CsvWriter::write (csv_writer.cpp:21-93), host only (no
 * device needed). append = 0 truncates the file and writes the header line;
 * append != 0 adds the rows of another subcase. Node rows go out in ascending
 * node ID order; element rows CQUAD4 first, then solids, each in element
 * order; values as "%.8e". Errors: VBH_ERR_SOLVER "Cannot write CSV: <path>". */
int vbh_write_node_csv(const char* path, int append, int subcase_id, int nnodes,
                       const int* node_ids, const double* disp);
int vbh_write_elem_csv(const char* path, int append, int subcase_id, int nelem,
                       const int* elem_ids, const int* etype, const double* rows);

/* ---- BDF input: BdfParser::parse_file (src/io/bdf_parser.cpp:78-633) for
 * the cards of this path (GRID, MAT1, PSOLID, PSHELL, CHEXA, CPENTA, CTETRA,
 * CQUAD4, FORCE, MOMENT, PLOAD4, GRAV, SPC, SPC1, MPC, MPCADD, RBE2, RBE3, CORD1R/C,
 * CORD2R/C/S, PARAM SHELLFORM; case control SUBCASE, LOAD, SPC, MPC,
 * DISPLACEMENT, STRESS) in small, large and free field with continuations.
 * Host only, no device needed; parsed in parallel chunks. Nodes come out in
 * ascending ID order (node indices below are 0-based into that order) with
 * positions in the basic frame (GRID CP resolved, Model::resolve_coordinates,
 * src/core/model.cpp:175-250), elements in file order, FORCE / MOMENT in a CID
 * frame rotated into basic. Higher-order solids, PLOAD4 direction vectors and
 * GRAV in a CID frame are rejected with VBH_ERR_SOLVER "Line N: ..."; other
 * cards are skipped and listed by vbh_deck_unsupported. */
typedef struct vbh_deck vbh_deck;
typedef struct {
  int nnodes, nelem, nconn, nsubcase;
  int solid_form;  /* VBH_SOLID_EAS / VBH_SOLID_SRI (PSOLID ISOP) */
  int shell_form;  /* VBH_SHELL_MITC4 / VBH_SHELL_MINDLIN (PARAM SHELLFORM) */
  int nunsupported;
  int nrigid_mpc, nrigid_terms;  /* RBE2 / RBE3 expanded into MPC equations */
} vbh_deck_sizes;
typedef struct {
  int id, load_set, spc_set;
  int nspc;          /* constrained nodes (masks OR-ed per node) */
  int npload4;       /* PLOAD4 faces of the load set */
  int disp_print, stress_print;  /* DISPLACEMENT / STRESS requests */
  double grav[3];    /* GRAV of the load set: scale x direction, basic frame */
  int mpc_set;       /* case control MPC = (0 = none) */
  int nmpc, nmpc_terms;  /* MPC equations of that set and their terms */
} vbh_deck_subcase_info;
int vbh_deck_read(const char* path, vbh_deck** out);
int vbh_deck_parse(const char* text, long long len, vbh_deck** out);
void vbh_deck_free(vbh_deck* deck);
int vbh_deck_info(const vbh_deck* deck, vbh_deck_sizes* out);
/* Any output pointer may be NULL. xyz nnodes x 3; eptr nelem + 1; E, nu, rho
 * from the element's MAT1 (NU blank = 0), thickness = PSHELL T (solids 0). */
int vbh_deck_mesh(const vbh_deck* deck, int* node_ids, double* xyz, int* elem_ids, int* etype,
                  int* eptr, int* conn, double* E, double* nu, double* rho, double* thickness);
int vbh_deck_subcase(const vbh_deck* deck, int k, vbh_deck_subcase_info* out);
/* Subcase k: SPC nodes and DOF masks (bit d = DOF d+1), nodal loads nnodes x 6
 * (FORCE into T1-T3, MOMENT into R1-R3), PLOAD4 element indices, face grids
 * G1 / G3-G4 as node indices (-1 = none) and 4 corner pressures each. */
int vbh_deck_subcase_data(const vbh_deck* deck, int k, int* spc_nodes, int* spc_masks,
                          double* nodal, int* pl_elem, int* pl_g1, int* pl_g34, double* pl_p);
/* MPC equations of subcase k's MPC set in model order (MPC cards in file
 * order, MPCADD copies appended where the MPCADD card stands): mpc_ptr
 * nmpc + 1 offsets into term_node (node index), term_dof (1..6), term_coeff. */
int vbh_deck_subcase_mpc(const vbh_deck* deck, int k, int* mpc_ptr, int* term_node, int* term_dof,
                         double* term_coeff);
/* RBE2 / RBE3 elements as MPC equations (expand_rbe2 / expand_rbe3,
 * src/elements/rbe_constraints.cpp:22-181; RBE2s then RBE3s, file order),
 * applied in every subcase ahead of the MPC set: mpc_ptr nrigid_mpc + 1,
 * terms as in vbh_deck_subcase_mpc; active[nnodes] the DOF mask (bit d =
 * DOF d+1) the rigid elements activate (mark_rigid_element_dofs,
 * src/solver/analysis_support.cpp:70-95). */
int vbh_deck_rigid(const vbh_deck* deck, int* active, int* mpc_ptr, int* term_node, int* term_dof,
                   double* term_coeff);
/* Displacement frames: cd[nnodes] the GRID CD (0 = basic) and rot[nnodes x 9]
 * the frame's rotation matrix at the node, row-major, columns the local unit
 * vectors in basic (identity for CD = 0). Node positions are already basic. */
int vbh_deck_frames(const vbh_deck* deck, int* cd, double* rot);
/* Keyword of the i-th unsupported bulk card, NULL past the end. */
const char* vbh_deck_unsupported(const vbh_deck* deck, int i);

/* ---- solve: SolverBackend::solve(K, F) -> u (solver_backend.hpp:26-29) ----
 * F and u are host arrays of length n. iters / rel_res may be NULL. */
int vbh_solve(vbh_ctx* ctx, const double* F, double* u, int* iters, double* rel_res);
/* Same with device-resident F and u (HBM pointers); no host copies. */
int vbh_solve_device(vbh_ctx* ctx, const double* dF, double* du, int* iters, double* rel_res);
/* vbh_solve_device without any device allocation. No reference
 * counterpart: it is for callers that cannot allocate device memory during a
 * solve (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments, results and errors as
 * vbh_solve_device on the operator set up before (vbh_set_csr /
 * vbh_assemble, which may allocate); the solver's work buffers live in one
 * static device array of 96 MiB per GPU, taken at vbh_ctx_create, and the
 * iterations are launched directly (no graph). Operators of more than
 * about 2,510,000 rows in FP64 (2,788,000 in mixed precision: 40 / 36 bytes
 * per row after 257 KiB of state and partial sums) are rejected with
 * VBH_ERR_ARG. Every call on one GPU shares that
 * GPU's array and keeps the last call's contents: calls that overlap on one
 * GPU (other threads or contexts) race. */
int vbh_solve_device_noalloc(vbh_ctx* ctx, const double* dF, double* du, int* iters,
                             double* rel_res);
/* Many load cases sharing K in one fused block-PCG (no reference counterpart:
 * the reference re-assembles and re-solves per SUBCASE, linear_static.cpp:202-320).
 * F and U are column-major n x k host arrays; iters/rel_res have length k. */
int vbh_solve_block(vbh_ctx* ctx, int k, const double* F, double* U, int* iters,
                    double* rel_res);
int vbh_solve_block_device(vbh_ctx* ctx, int k, const double* dF, double* dU, int* iters,
                           double* rel_res);
/* SpMM configuration the block solver's autotuner chose for the last block
 * solve (no reference counterpart): column tile (load cases per pass over K),
 * load cases per lane, lanes per block-row group (64 / lanes_per_group block
 * rows per wave), grid rows, whether block rows are split into one
 * contiguous range per XCD, how each lane gets the block values
 * (value_path: 0 direct loads, 1 staged through LDS, 2 broadcast across
 * the 16-lane DPP row), padded column count and the measured mean SpMM
 * launch in ms. */
typedef struct {
  int col_tile;
  int cols_per_lane;
  int lanes_per_group;
  int row_blocks;
  int xcd_rows;
  int value_path;
  int ldk;
  double spmm_ms;
} vbh_block_config;
int vbh_last_block_config(const vbh_ctx* ctx, vbh_block_config* out);
int vbh_last_stats(const vbh_ctx* ctx, vbh_solve_stats* out);

/* ---- operator application ------------------------------------------------
 * y = K x in FP64 with the context's device operator (reference
 * CsrData::multiply, sparse_matrix.cpp:152-181). Host arrays of length n. */
int vbh_spmv(vbh_ctx* ctx, const double* x, double* y);
/* Time `reps` back-to-back launches of one solver kernel on device-resident
 * vectors; returns the mean launch duration in ms measured with hipEvents on
 * the context stream. which: 0/1 plain SpMV FP64/FP32, 2/3 the PCG SpMV
 * kernel (K p + the p.Ap, z.Ap, Ap.diag^-1.Ap partial sums) FP64/FP32, 4/5
 * the PCG vector kernel (update + next search direction) FP64/FP32, 6 a
 * 256 MiB device-to-device streaming copy (calibration of byte counters);
 * 7..10 time one kernel inside `reps` directly launched PCG iterations
 * (SpMV, vector kernel) from the start and stop timestamps of that kernel's
 * own dispatch (hipExtLaunchKernelGGL events): 7/8 the SpMV FP64/FP32, 9/10
 * the vector kernel; 13/14 the SpMV FP64/FP32 launched `reps` times back to
 * back between one event pair (kernel plus dispatch gap, no per-launch event
 * cost). 11 and 12 are not used. 2..5 and 7..14 use the F of the last
 * solve. */
int vbh_bench_spmv(vbh_ctx* ctx, int which, int reps, double* ms_per_launch);
/* Bytes one launch of the single-load-case solver SpMV reads from the device
 * operator in the layout it actually streams (no reference counterpart;
 * measurement): precision VBH_PREC_FP64 or VBH_PREC_MIXED (the FP32 stream).
 * Sliced ELL: padded block values, block columns (16- or 32-bit), slice
 * offsets and, for the split FP32 stream, the FP64 diagonal node blocks;
 * scalar CSR: values, column indices and row pointers. Vector streams (p
 * gathered, Ap written) are not included. */
int vbh_operator_stream_bytes(const vbh_ctx* ctx, int precision, long long* bytes);

/* ---- device memory helpers for device-resident workflows ---------------- */
int vbh_device_alloc(vbh_ctx* ctx, long long bytes, void** dptr);
int vbh_device_free(vbh_ctx* ctx, void* dptr);
int vbh_memcpy_h2d(vbh_ctx* ctx, void* dst, const void* src, long long bytes);
int vbh_memcpy_d2h(vbh_ctx* ctx, void* dst, const void* src, long long bytes);

#ifdef __cplusplus
}
#endif
#endif /* VBH_H */
