// pstm_fwd_f32.hip -- forward tile kernels for PSTM_F32 traces (one
// translation unit per trace dtype so the instantiations compile in parallel).
#include "pstm_tile.h"

namespace pstm {

template <>
int launch_forward_dt<PSTM_F32>(const pstm_params_t* p, const PlanLayout& L, const TileArgs& a,
                               hipStream_t s) {
  return launch_forward_impl<PSTM_F32>(p, L, a, s);
}

}  // namespace pstm
