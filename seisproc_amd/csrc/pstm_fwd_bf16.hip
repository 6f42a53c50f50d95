// pstm_fwd_bf16.hip -- forward tile kernels for PSTM_BF16 traces (one
// translation unit per trace dtype so the instantiations compile in parallel).
#include "pstm_tile.h"

namespace pstm {

template <>
int launch_forward_dt<PSTM_BF16>(const pstm_params_t* p, const PlanLayout& L, const TileArgs& a,
                               hipStream_t s) {
  return launch_forward_impl<PSTM_BF16>(p, L, a, s);
}

}  // namespace pstm
