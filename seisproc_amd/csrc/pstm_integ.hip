// pstm_integ.hip -- causal and anticausal integrals of whole traces, the
// inputs of the antialias forward when the caller integrates once per call
// instead of once per staged window.
//
// The triangle antialias filter of half-width L at position u is
// [2 A(u) - A(u - L) - A(u + L)] / L^2 with A the anticausal integral of the
// causal integral C of the trace (zero outside the record).  A second
// difference is unchanged by a constant or linear term, so integrating the
// whole trace here gives the same taps as integrating each window in LDS:
// the windows are then copies (with the record's end extended linearly).
#include "pstm_tile.h"

namespace pstm {

// One wave per trace: each lane sums a contiguous segment of samples, a wave
// scan joins the segments (causal pass), then the same for the anticausal
// pass over C, which A holds in FP64 between the passes.
template <int DT>
__global__ __launch_bounds__(kThreads) void k_integrate(const void* traces, const float* scale, int64_t n_traces,
                                                        int n, double* A, float* C) {
  const int lane = threadIdx.x & 63;
  const int64_t t = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);  // uniform per wave
  if (t >= n_traces) return;
  const float sc = scale ? scale[t] : 1.f;
  const int64_t row = t * n;
  const int seg = (n + 63) >> 6;
  const int s0 = min(lane * seg, n), s1 = min(s0 + seg, n);
  double tot = 0.0;
  for (int s = s0; s < s1; ++s) tot += (double)(load_sample<DT>(traces, row + s) * sc);
  double run = wave_scan(tot, lane) - tot;  // the samples before this lane's segment
  double csum = 0.0;
  for (int s = s0; s < s1; ++s) {
    run += (double)(load_sample<DT>(traces, row + s) * sc);
    A[row + s] = run;
    C[row + s] = (float)run;
    csum += run;
  }
  const double incl = wave_scan(csum, lane);
  double after = lane63(incl) - incl;  // C summed over the later segments
  for (int s = s1 - 1; s >= s0; --s) {
    after += A[row + s];
    A[row + s] = after;
  }
}

int launch_integrate(const pstm_params_t* p, const void* traces, const float* scale, double* int_a, float* int_c,
                     hipStream_t s) {
  const int64_t blocks = (p->n_traces + kWaves - 1) / kWaves;
  if (blocks <= 0) return PSTM_OK;
  if (blocks > 0x7fffffff) return set_error(PSTM_ERR_INVALID, "too many traces (%lld)", (long long)p->n_traces);
  if (p->dtype == PSTM_F32)
    hipLaunchKernelGGL(k_integrate<PSTM_F32>, dim3((unsigned)blocks), dim3(kThreads), 0, s, traces, scale,
                       (int64_t)p->n_traces, p->n_samples, int_a, int_c);
  else if (p->dtype == PSTM_BF16)
    hipLaunchKernelGGL(k_integrate<PSTM_BF16>, dim3((unsigned)blocks), dim3(kThreads), 0, s, traces, scale,
                       (int64_t)p->n_traces, p->n_samples, int_a, int_c);
  else
    hipLaunchKernelGGL(k_integrate<PSTM_F16>, dim3((unsigned)blocks), dim3(kThreads), 0, s, traces, scale,
                       (int64_t)p->n_traces, p->n_samples, int_a, int_c);
  return check_launch("integrate");
}

}  // namespace pstm
