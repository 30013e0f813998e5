// vibestran_amd/csrc/vbh_ctx.hpp -- the context behind vbh_ctx*.
#pragma once
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "vbh_internal.hpp"

namespace vbh {

struct BlockState;  // block-PCG work space (vbh_block.hip)
void destroy_block_state(BlockState* b);

struct GraphSlot {
  hipGraphExec_t exec = nullptr;
  int iters = 0;
  double tol = 0.0;
  int max_iters = 0;
};

// Multi-point constraint elimination on the host (vbh_mpc.hip): the reduced
// numbering and T (n_full x n_red, CSR by full rows) with its transpose.
struct MpcSystem {
  bool active = false;
  int n_full = 0, n_red = 0, n_elim = 0, n_skipped = 0;
  std::vector<int> index_map;                // full -> reduced equation, -1 dependent
  std::vector<int> red_to_full;              // reduced -> full equation
  std::vector<int> deps;                     // dependent full equations, ascending
  std::vector<int> t_ptr, t_col, tt_ptr, tt_row;
  std::vector<double> t_val, tt_val;
  void build(int nnodes, const int* eq, int n_full, int nmpc, const int* ptr, const int* node,
             const int* dof, const double* coeff);
};

// T of the mesh's elimination on the device (inactive: identity, n_red = n_full).
struct MpcDev {
  bool active = false;
  int n_full = 0, n_red = 0, n_dep = 0;
  DBuf<int> t_ptr, t_col, tt_ptr, tt_row;
  DBuf<int> t_colF, deps;                    // T's columns as full equations; dependents
  DBuf<double> t_val, tt_val;
  void upload(const MpcSystem& h, int n_mesh, hipStream_t s);
  void expand(const double* u_red, double* u_full, hipStream_t s) const;   // T u
  void reduce(const double* f_full, double* f_red, hipStream_t s) const;   // T^T f
};

// The mesh of the last vbh_assemble, kept on the device for the element
// sensitivity kernels (and anything else that revisits the elements). n is
// the size of the mesh's equation numbering (eq); under MPCs the operator has
// mpc.n_red equations.
struct MeshDev {
  bool valid = false;
  int n = 0, nnodes = 0, nelem = 0, solid_form = 0, shell_form = 0;
  DBuf<double> xyz, E, nu, t;
  DBuf<int> eq, et, ep, conn, gd;
  DBuf<int> sel[5];
  int nsel[5] = {0, 0, 0, 0, 0};
  MpcDev mpc;
};

struct Ctx {
  // device / stream
  int device = 0;
  int n_cu = 256;  // compute units of the device
  std::string name, arch;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  hipEvent_t ev_poll[2] = {nullptr, nullptr};  // state copies of the last two iteration batches

  // solver configuration (CudaPCGContext equivalent, impl:2030-2040)
  double tol = 1e-8;
  int max_iters = 10000;
  int precision = VBH_PREC_FP64;
  int format_req = -1;
  // mixed precision: FP64 refresh once the recursive residual fell by this
  // factor; 0 = by operator: 0.2 for solids (and scalar CSR), 0.8 for 6x6
  // shell blocks, whose iteration counts gain from frequent refreshes (C3
  // block solve 20.5k mean iterations at 0.8 against 23.0k at 0.2; a solid
  // block loses 5 % at 0.4)
  double refresh_delta = 0.0;
  double delta() const { return refresh_delta > 0.0 ? refresh_delta : (format == 6 ? 0.8 : 0.2); }
  int spmv_acc64 = 0;  // FP32 SpMV rows accumulated in FP64
  int iters_per_graph = 32;

  // operator
  bool have_matrix = false;
  int n = 0, n_pad = 0, format = 0;
  // Under an MPC elimination the operator keeps the full equation numbering
  // (dependent rows and columns an identity block, so node blocks stay
  // aligned) while the API's vectors and CSR are the reduced system: n_api
  // equations, op_api_of[i] the API index of operator row i (-1: dependent),
  // op_full_of[j] the operator row of API index j.
  bool op_padded = false;
  int n_api = 0;
  DBuf<int> op_api_of, op_full_of;
  std::vector<int> op_api_host;  // op_api_of on the host
  long long operator_version = 0;
  std::vector<int> lower_rp, lower_ci;    // reference CsrData (lower) on the host
  std::vector<double> lower_v;
  bool lower_on_host = false;
  int l_nnz = 0;
  int block_hint = 3;
  DBuf<int> l_rp, l_ci, l_row;            // device lower CSR (assembled operators)
  DBuf<double> l_v;
  vbh_assembly_stats asm_stats{};
  DBuf<int> csr_rp, csr_ci;
  DBuf<double> csr_v64;
  DBuf<float> csr_v32;
  DBuf<int> bsr_rp, bsr_ci;
  DBuf<double> bsr_v64;
  DBuf<float> bsr_v32;
  DBuf<int> sell_koff, sell_idx;
  DBuf<short> sell_idx16;
  DBuf<double> sell_v64;
  DBuf<float> sell_v32;
  DBuf<double> sell_dblk;   // FP64 corrected diagonal node blocks (row-sum preserving FP32)
  DBuf<int> sell_kdiag;     // slot of the diagonal block in each block row
  DBuf<int> sell_lrows;     // block rows longer than kSellCap blocks
  int rowsum_fp32 = 1;      // split the FP32 stream as above (VBH_ROWSUM=0 disables)
  long long sell_padded_blocks = 0;
  long long sell_slots = 0;  // block slots of all slices (sum of the slice widths)
  // relative sliced ELL (BsrDev::rk64 ...)
  DBuf<int> rel_k64, rel_o64, rel_k32, rel_o32;
  DBuf<double> rel_v64;
  DBuf<float> rel_v32;
  long long rel_slots64 = 0, rel_slots32 = 0;
  DBuf<double> diag64;
  DBuf<float> diag32;
  CsrDev csr;
  BsrDev bsrd;

  // solve work space
  size_t work_len = 0;
  long long work_version = -1;
  DBuf<double> w_F, w_u, w_r64, w_r, w_p, w_Ap;
  DBuf<float> w_r32, w_p32, w_Ap32;
  DBuf<double> part_a, part_b;
  int red_n_a = 1, red_n_b = 1;
  DBuf<PcgScalars> st_dev;
  PcgScalars* st_host = nullptr;  // pinned, two slots (see drive)
  GraphSlot graphs[2];

  // mesh of the last device assembly
  MeshDev mesh;
  // MPC elimination applied by the next vbh_assemble (vbh_set_mpc)
  MpcSystem mpc;

  // block PCG over several load cases
  std::unique_ptr<BlockState, void (*)(BlockState*)> blk{nullptr, destroy_block_state};

  // diagnostics (SolverBackend::last_iteration_count / last_estimated_error)
  int last_iters = 0;
  double last_rel = 0.0;
  vbh_solve_stats stats{};

  ~Ctx() {
    if (device >= 0) (void)hipSetDevice(device);
    drop_graphs();
    if (st_host) (void)hipHostFree(st_host);
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_stop) (void)hipEventDestroy(ev_stop);
    for (hipEvent_t e : ev_poll)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void select_stream_device();
  void upload_operator(int n, int nnz, const int* rp, const int* ci, const double* v, bool lower);
  void ensure_work(int nvec_len);
  void assemble(int n, int nnodes, const double* xyz, const int* eq, int nelem, const int* etype,
                const int* eptr, const int* conn, const double* E, const double* nu,
                const double* thick, int solid_form, int shell_form);
  void finish_device_operator();
  void finish_block_operator(int B);
  bool set_lower_device(int n, int nnz, const int* rp, const int* ci, const double* v);
  void apply_mpc_to_lower();
  // API vector (n_api, host or device) <-> operator vector (n, device)
  void to_operator(const double* x, bool dev, double* dst, int k = 1);
  void from_operator(const double* src, double* x, bool dev, int k = 1);
  const double* mesh_solution(const double* u, bool dev, DBuf<double>& up, DBuf<double>& full);
  void element_sensitivity(int param, const double* u, const double* v, double* out, bool dev);
  // full: F in node-DOF layout (6 per node, every DOF) instead of equations
  void assemble_rhs(const double* nodal, int nface, const int* face_nodes, const double* face_p,
                    const double* grav, const double* rho, double* F, bool dev,
                    bool full = false);
  void recover_stress(const double* u, double* centroid, double* corner, bool dev);
  void element_forces(const double* u, double* fout, bool dev);
  void reactions(const double* u, const double* nodal, int nface, const int* face_nodes,
                 const double* face_p, const double* grav, const double* rho, double* R);
  void recover_displacements(const double* u, double* disp, bool dev);
  void fetch_lower_to_host();
  int dof_map(int nnodes, const int* ids, int nelem, const int* etype, const int* eptr,
              const int* conn, int nspc, const int* spc_nodes, const int* spc_masks, int* eq_out);
  void drop_graphs();
  PcgParams params() const;
  CgRed red() const;
  CgBufs<double> bufs64();
  CgBufs<float> bufs32();
  void enqueue_iterations(int precision_mode, int count);
  hipGraphExec_t iteration_graph(int precision_mode);
  PcgScalars drive(const PcgScalars* st, const std::function<void()>& batch);
  PcgScalars run_solve(int precision_mode);
  void check_outcome(const PcgScalars& s, int precision_mode);
};

}  // namespace vbh
