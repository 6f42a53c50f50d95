// pstm_static.hip -- pstm_migrate_static: plan build and forward migration
// in one call that allocates nothing, for callers that cannot call hipMalloc
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// allocation).  The plan workspace the other entry points take from the
// caller lives here in a file-scope device array of fixed size.
#include <mutex>

#include "pstm_internal.h"

namespace pstm {

// Plan workspace of pstm_migrate_static: 32 MiB per GPU, held from the
// library's load on a device until the process ends.  In host code the name
// is only a host-side stand-in: its device address comes from
// hipGetSymbolAddress, once per device.
constexpr size_t kStaticScratchBytes = PSTM_STATIC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices];
static std::once_flag g_scratch_once[kMaxDevices];
static hipError_t g_scratch_err[kMaxDevices];

// Device address of `scratch` on the current device (set up on the first
// call on that device, with that device current).
static int scratch_on_current_device(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_hip_error("hipGetDevice", e);
  if (dev < 0 || dev >= kMaxDevices) return set_error(PSTM_ERR_UNSUPPORTED, "device ordinal %d >= %d", dev, kMaxDevices);
  std::call_once(g_scratch_once[dev], [dev]() {
    void* p = nullptr;
    g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? p : nullptr;
  });
  if (g_scratch_err[dev] != hipSuccess) return set_hip_error("hipGetSymbolAddress(scratch)", g_scratch_err[dev]);
  *out = g_scratch_ptr[dev];
  return PSTM_OK;
}

}  // namespace pstm

using namespace pstm;

extern "C" {

int pstm_migrate_static(const pstm_params_t* p, const double* src_xy, const double* rcv_xy,
                        const void* traces, const float* trace_scale, const float* vel, const float* eta,
                        float* image, float* image_comp, float* fold, void* stream) {
  const size_t need = pstm_plan_bytes(p);  // validates p, sets the message on failure
  if (need == 0) return PSTM_ERR_INVALID;
  if (need > kStaticScratchBytes)
    return set_error(PSTM_ERR_WORKSPACE,
                     "pstm_migrate_static: the plan of %lld traces needs %zu bytes, above the static "
                     "workspace of %zu bytes; split the traces over several calls or use pstm_plan_build",
                     (long long)p->n_traces, need, kStaticScratchBytes);
  void* ws = nullptr;
  int rc = scratch_on_current_device(&ws);
  if (rc != PSTM_OK) return rc;
  rc = pstm_plan_build(p, src_xy, rcv_xy, ws, kStaticScratchBytes, stream);
  if (rc != PSTM_OK) return rc;
  return pstm_forward(p, ws, traces, trace_scale, vel, eta, image, image_comp, fold, stream);
}

}  // extern "C"
