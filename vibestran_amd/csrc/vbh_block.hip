// vibestran_amd/csrc/vbh_block.hip -- block PCG over k load cases (placeholder
// until the fused block kernels land: reports VBH_ERR_STATE).
#include "vbh_ctx.hpp"

extern "C" {
int vbh_solve_block(vbh_ctx*, int, const double*, double*, int*, double*) {
  vbh::set_error("vbh_solve_block: block PCG not available in this build");
  return VBH_ERR_STATE;
}
int vbh_solve_block_device(vbh_ctx*, int, const double*, double*, int*, double*) {
  vbh::set_error("vbh_solve_block_device: block PCG not available in this build");
  return VBH_ERR_STATE;
}
}
