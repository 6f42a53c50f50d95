// pstm_grad.hip -- velocity-gradient tile kernels (isotropic, no antialias).
#include "pstm_tile.h"

namespace pstm {

// 3 workgroups per CU (168 VGPRs, a small spill): 28% faster than 2 on a
// 4-class slice of the LSM workload, 4 spills more and is slower again.  The
// common case (FULL weight, no taper, clamped record end) compiles the weight
// mode in (GEN = false): 2728 ms against 3598 for the general form on that
// slice, with the same gradient.
int launch_vel_grad(const pstm_params_t* p, const PlanLayout& L, const TileArgs& a, hipStream_t s) {
  int rc;
  if (long_traces(p, a)) {
    // windows may exceed the LDS share: the general form staging such
    // windows in segments
#define PSTM_GRADS(TX, TY, DTV) launch_one<MODE_GRAD, TX, TY, DTV, false, false, true, 3, false, false, PSTM_WEIGHT_FULL, true>(a, s)
#define PSTM_GRADS3(TX, TY)                                                     \
  (p->dtype == PSTM_F32 ? PSTM_GRADS(TX, TY, PSTM_F32)                          \
                        : p->dtype == PSTM_BF16 ? PSTM_GRADS(TX, TY, PSTM_BF16) \
                                                : PSTM_GRADS(TX, TY, PSTM_F16))
    if (a.tile_x == 4 && a.tile_y == 2) rc = PSTM_GRADS3(4, 2);
    else if (a.tile_x == 8 && a.tile_y == 1) rc = PSTM_GRADS3(8, 1);
    else if (a.tile_x == 1 && a.tile_y == 8) rc = PSTM_GRADS3(1, 8);
    else rc = set_error(PSTM_ERR_INVALID, "unsupported tile shape %dx%d", a.tile_x, a.tile_y);
#undef PSTM_GRADS3
#undef PSTM_GRADS
    return rc != PSTM_OK ? rc : check_launch("vel_grad");
  }
  const bool gen = !fast_path(p);
#define PSTM_GRAD(TX, TY, DTV, GENV) launch_one<MODE_GRAD, TX, TY, DTV, false, false, GENV, 3>(a, s)
#define PSTM_GRAD3(TX, TY, GENV)                                                      \
  (p->dtype == PSTM_F32 ? PSTM_GRAD(TX, TY, PSTM_F32, GENV)                           \
                        : p->dtype == PSTM_BF16 ? PSTM_GRAD(TX, TY, PSTM_BF16, GENV)   \
                                                : PSTM_GRAD(TX, TY, PSTM_F16, GENV))
  if (a.tile_x == 4 && a.tile_y == 2) rc = gen ? PSTM_GRAD3(4, 2, true) : PSTM_GRAD3(4, 2, false);
  else if (a.tile_x == 8 && a.tile_y == 1) rc = PSTM_GRAD3(8, 1, true);
  else if (a.tile_x == 1 && a.tile_y == 8) rc = PSTM_GRAD3(1, 8, true);
  else rc = set_error(PSTM_ERR_INVALID, "unsupported tile shape %dx%d", a.tile_x, a.tile_y);
#undef PSTM_GRAD3
#undef PSTM_GRAD
  return rc != PSTM_OK ? rc : check_launch("vel_grad");
}

}  // namespace pstm
