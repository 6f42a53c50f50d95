// valu_rates.hip -- issue rates of the VALU instruction kinds the tile
// kernels are built from, on the GPU this runs on.  A standalone probe (not
// part of libpstm_hip.so); its figures bound the tile kernels' roofline:
//
//   ./seisproc_amd/valu_rates            one JSON line per instruction kind
//
// Each kernel runs 8 independent dependency chains per lane (8 waves per SIMD,
// so issue, not latency, limits it) of one instruction kind, and reports lane
// operations per second.  The mixes replay the forward inner loop's 13
// instructions per contribution, and a square-root / reciprocal form of it.
// Compiled like the library (no packed FP32): the v_fma_f32 line is the
// unpacked FP32 rate the kernels can use.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                       \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(1);                                                                    \
    }                                                                                  \
  } while (0)

namespace {

constexpr int kChains = 8;

enum Op { FMA, MUL, ADD, RSQ, SQRT, RCP, FRACT, CVT, LSHL_ADD, ADD_U32, MED3, MIX_FWD, MIX_SQRT };

template <int OP>
__global__ __launch_bounds__(256) void k_rate(float* out, int iters, float b, float c) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  float x[kChains];
  int xi[kChains];
#pragma unroll
  for (int i = 0; i < kChains; ++i) {
    x[i] = 1.0f + 1e-3f * (float)(gid + i);
    xi[i] = gid + i;
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < kChains; ++i) {
      if constexpr (OP == FMA) x[i] = fmaf(x[i], b, c);
      if constexpr (OP == MUL) x[i] = x[i] * b;
      if constexpr (OP == ADD) x[i] = x[i] + b;
      if constexpr (OP == RSQ) x[i] = __builtin_amdgcn_rsqf(x[i]);
      if constexpr (OP == SQRT) x[i] = __builtin_amdgcn_sqrtf(x[i]);
      if constexpr (OP == RCP) x[i] = __builtin_amdgcn_rcpf(x[i]);
      if constexpr (OP == FRACT) x[i] = __builtin_amdgcn_fractf(x[i]);
      if constexpr (OP == CVT) xi[i] = (int)__int_as_float(xi[i] | 0x3f800000);  // v_or + v_cvt_i32_f32
      if constexpr (OP == LSHL_ADD) xi[i] = (xi[i] << 3) + (int)c;
      if constexpr (OP == ADD_U32) xi[i] = xi[i] + (int)c;
      if constexpr (OP == MED3) x[i] = __builtin_amdgcn_fmed3f(x[i], b, c);
      if constexpr (OP == MIX_FWD) {
        // forward inner loop: 2 FMA, 2 v_rsq, 2 FMA, v_cvt, v_fract,
        // v_lshl_add, lerp FMA, 2 MUL, accumulate FMA (the LDS read aside)
        const float xs = fmaf(x[i], b, c), xr = fmaf(x[i], c, b);
        const float rs = __builtin_amdgcn_rsqf(xs), rr = __builtin_amdgcn_rsqf(xr);
        const float u = fmaf(xs, rs, fmaf(xr, rr, -c));
        const int i0 = (int)u;
        const float f = __builtin_amdgcn_fractf(u);
        const int ad = (i0 << 3) + xi[i];
        const float p = rs * rr, w = p * p;
        x[i] = fmaf(w, fmaf(f, __int_as_float(ad), b), x[i]);
      }
      if constexpr (OP == MIX_SQRT) {
        // the same with the legs as v_sqrt and the weight as v_rcp of their product
        const float xs = fmaf(x[i], b, c), xr = fmaf(x[i], c, b);
        const float u = __builtin_amdgcn_sqrtf(xs) + __builtin_amdgcn_sqrtf(xr);
        const int i0 = (int)u;
        const float f = __builtin_amdgcn_fractf(u);
        const int ad = (i0 << 3) + xi[i];
        const float w = __builtin_amdgcn_rcpf(xs * xr);
        x[i] = fmaf(w, fmaf(f, __int_as_float(ad), b), x[i]);
      }
    }
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < kChains; ++i) s += x[i] + (float)xi[i];
  out[gid] = s;  // keeps every chain live
}

template <int OP>
void run(const char* name, const char* insts, int ops, float* d, int blocks, int iters, double fma_rate,
         double* rate_out) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  k_rate<OP><<<blocks, 256>>>(d, 16, 0.999f, 0.001f);
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  float best = 1e30f;
  for (int rep = 0; rep < 3; ++rep) {
    CHECK(hipEventRecord(e0));
    k_rate<OP><<<blocks, 256>>>(d, iters, 0.999f, 0.001f);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    best = ms < best ? ms : best;
  }
  const double lane_ops = (double)blocks * 256 * iters * kChains * ops;
  const double rate = lane_ops / (best * 1e-3) / 1e12;  // T lane-op/s
  if (rate_out) *rate_out = rate;
  const double fma_units = fma_rate > 0 ? fma_rate * ops / rate : 0.0;  // time per chain step in FMA times
  std::printf("{\"kind\": \"%s\", \"instructions\": \"%s\", \"ops_per_step\": %d, \"ms\": %.3f, "
              "\"T_lane_op_s\": %.2f, \"fma_times_per_step\": %.2f}\n",
              name, insts, ops, best, rate, fma_units);
  std::fflush(stdout);
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
}

}  // namespace

int main() {
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  std::printf("{\"device\": \"%s\", \"cus\": %d, \"clock_khz\": %d}\n", prop.gcnArchName, prop.multiProcessorCount,
              prop.clockRate);
  const int blocks = prop.multiProcessorCount * 8;  // 8 workgroups of 4 waves per CU: 8 waves per SIMD
  float* d = nullptr;
  CHECK(hipMalloc(&d, (size_t)blocks * 256 * sizeof(float)));
  const int it = 20000;
  double fma = 0.0;
  run<FMA>("fma_f32", "v_fma_f32", 1, d, blocks, it, 0.0, &fma);
  run<MUL>("mul_f32", "v_mul_f32", 1, d, blocks, it, fma, nullptr);
  run<ADD>("add_f32", "v_add_f32", 1, d, blocks, it, fma, nullptr);
  run<RSQ>("rsq_f32", "v_rsq_f32", 1, d, blocks, it, fma, nullptr);
  run<SQRT>("sqrt_f32", "v_sqrt_f32", 1, d, blocks, it, fma, nullptr);
  run<RCP>("rcp_f32", "v_rcp_f32", 1, d, blocks, it, fma, nullptr);
  run<FRACT>("fract_f32", "v_fract_f32", 1, d, blocks, it, fma, nullptr);
  run<CVT>("or+cvt_i32", "v_or_b32 + v_cvt_i32_f32", 2, d, blocks, it, fma, nullptr);
  run<LSHL_ADD>("lshl_add_u32", "v_lshl_add_u32", 1, d, blocks, it, fma, nullptr);
  run<ADD_U32>("add_u32", "v_add_u32", 1, d, blocks, it, fma, nullptr);
  run<MED3>("med3_f32", "v_med3_f32", 1, d, blocks, it, fma, nullptr);
  run<MIX_FWD>("forward_mix", "2 fma, 2 rsq, 2 fma, cvt, fract, lshl_add, fma, 2 mul, fma", 13, d, blocks, it / 4,
               fma, nullptr);
  run<MIX_SQRT>("sqrt_rcp_mix", "2 fma, 2 sqrt, add, cvt, fract, lshl_add, mul, rcp, 2 fma", 12, d, blocks, it / 4,
                fma, nullptr);
  CHECK(hipFree(d));
  return 0;
}
