// pstm_internal.h -- declarations shared by the library's translation units.
#pragma once

#include "pstm_common.h"

namespace pstm {

// Error reporting (thread-local message, see pstm_last_error()).
int set_error(int status, const char* fmt, ...);
int set_hip_error(const char* what, hipError_t e);
int check_launch(const char* what);

// Host-side layout of a plan for `p` (no GPU access).  Returns PSTM_OK or an
// error status with the message set.
int compute_layout(const pstm_params_t* p, PlanLayout* out);

size_t sort_temp_bytes(int64_t n, int key_bits);
int plan_build(const pstm_params_t* p, const PlanLayout& L, const double* src_xy,
               const double* rcv_xy, void* ws, hipStream_t s);
int quantize(const float* x, int64_t n_traces, int32_t n_samples, int32_t dtype, void* out,
             float* scale, hipStream_t s);

// Fill the kernel argument block from params + plan (no GPU access).
void fill_tile_args(const pstm_params_t* p, const PlanLayout& L, const void* ws, TileArgs* a);

int launch_forward(const pstm_params_t* p, const PlanLayout& L, const TileArgs& a, hipStream_t s);
// Causal and anticausal integrals of every trace (antialias forward inputs).
int launch_integrate(const pstm_params_t* p, const void* traces, const float* scale, double* int_a,
                     float* int_c, hipStream_t s);
int launch_adjoint(const pstm_params_t* p, const PlanLayout& L, const TileArgs& a, hipStream_t s);
int launch_vel_grad(const pstm_params_t* p, const PlanLayout& L, const TileArgs& a,
                    hipStream_t s);

}  // namespace pstm
