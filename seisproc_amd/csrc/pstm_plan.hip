// pstm_plan.hip -- trace binning for the tile kernels.
//
// Replaces the per-tile trace pre-filter of the reference
// (processors/migration/kirchhoff_kernel.py:579-590, a distance test of every
// trace midpoint against every 100-point tile) with a one-off binning: every
// trace is reduced to image-grid-local FP32 coordinates (FP64 arithmetic, once
// per trace), keyed by the spatial cell of its midpoint, and stably
// radix-sorted so that the traces of any rectangle of cells are a few
// contiguous index ranges.
#include <hipcub/hipcub.hpp>

#include "pstm_internal.h"

namespace pstm {

__global__ void k_plan_header(PlanHeader* h, int64_t n_traces, int32_t ncx, int32_t ncy) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    h->magic = kPlanMagic;
    h->n_traces = n_traces;
    h->n_kept = 0;
    h->ncx = ncx;
    h->ncy = ncy;
  }
}

// FP64 world -> FP32 local (along il, along xl) relative to the image origin;
// key = offset band * cells + midpoint cell.
__global__ void k_plan_local(const double* __restrict__ src, const double* __restrict__ rcv,
                             int64_t n, double ox, double oy, double ux_x, double ux_y,
                             double uy_x, double uy_y, double lo_a, double hi_a, double lo_b,
                             double hi_b, double cx0, double cy0, double cw, double ch,
                             int32_t ncx, int32_t ncy, int32_t n_bands, double two_a,
                             double omin, double omax, float4* __restrict__ local,
                             float* __restrict__ offs, uint32_t* __restrict__ keys,
                             int32_t* __restrict__ ids) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double sx = src[2 * j] - ox, sy = src[2 * j + 1] - oy;
  double rx = rcv[2 * j] - ox, ry = rcv[2 * j + 1] - oy;
  double sa = sx * ux_x + sy * ux_y, sb = sx * uy_x + sy * uy_y;
  double ra = rx * ux_x + ry * ux_y, rb = rx * uy_x + ry * uy_y;
  double ma = 0.5 * (sa + ra), mb = 0.5 * (sb + rb);
  const uint32_t n_cells = (uint32_t)ncx * (uint32_t)ncy;
  uint32_t key = n_cells * (uint32_t)n_bands;  // dropped
  // A trace can only contribute if its midpoint is within the aperture of
  // some image column (|m - x| <= (h_s + h_r)/2 < aperture) and its offset is
  // below twice the aperture (h_s + h_r >= offset).
  // With an offset window (omax > 0) only traces inside [omin, omax] are
  // kept (processors/migration/aperture.py:184-224, decided in FP64).
  const double off = hypot(sa - ra, sb - rb);
  const bool in_window = !(omax > 0) || (off >= omin && off <= omax);
  if (ma > lo_a && ma < hi_a && mb > lo_b && mb < hi_b && off < two_a && in_window) {
    int32_t cx = (int32_t)floor((ma - cx0) / cw);
    int32_t cy = (int32_t)floor((mb - cy0) / ch);
    cx = cx < 0 ? 0 : (cx >= ncx ? ncx - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= ncy ? ncy - 1 : cy);
    int32_t band = (int32_t)(off * (double)n_bands / two_a);
    band = band >= n_bands ? n_bands - 1 : band;
    key = (uint32_t)band * n_cells + (uint32_t)cy * (uint32_t)ncx + (uint32_t)cx;
  }
  local[j] = make_float4((float)sa, (float)sb, (float)ra, (float)rb);
  offs[j] = (float)off;
  keys[j] = key;
  ids[j] = (int32_t)j;
}

__global__ void k_plan_gather(const float4* __restrict__ local, const float* __restrict__ offs_in,
                              const int32_t* __restrict__ ids, int64_t n, float4* __restrict__ geom,
                              float* __restrict__ offs) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t j = ids[i];
  geom[i] = local[j];
  offs[i] = offs_in[j];
}

// cell_start[c] = first sorted index with key >= c, for c in [0, n_cells]
// (n_cells counts every (band, cell) key).
__global__ void k_plan_cells(const uint32_t* __restrict__ keys, int64_t n, uint32_t n_cells,
                             int32_t* __restrict__ cell_start, PlanHeader* h) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > (int64_t)n_cells) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < (uint32_t)c) lo = mid + 1; else hi = mid;
  }
  cell_start[c] = (int32_t)lo;
  if (c == (int64_t)n_cells) h->n_kept = lo;
}

size_t sort_temp_bytes(int64_t n, int key_bits) {
  size_t bytes = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                     (const int32_t*)nullptr, (int32_t*)nullptr, (int)n, 0,
                                     key_bits, (hipStream_t)0);
  return bytes;
}

int plan_build(const pstm_params_t* p, const PlanLayout& L, const double* src_xy,
               const double* rcv_xy, void* ws, hipStream_t s) {
  char* base = (char*)ws;
  auto* hdr = (PlanHeader*)(base + L.off_header);
  auto* keys_in = (uint32_t*)(base + L.off_keys_in);
  auto* keys_out = (uint32_t*)(base + L.off_keys_out);
  auto* ids_in = (int32_t*)(base + L.off_ids_in);
  auto* ids_out = (int32_t*)(base + L.off_ids_out);
  auto* local = (float4*)(base + L.off_local);
  auto* geom = (float4*)(base + L.off_geom);
  auto* cells = (int32_t*)(base + L.off_cell_start);
  auto* offs_local = (float*)(base + L.off_offs_local);
  auto* offs = (float*)(base + L.off_offs);
  const int64_t n = L.n_traces;
  const uint32_t n_cells = (uint32_t)L.ncx * (uint32_t)L.ncy * (uint32_t)L.n_bands;  // sort keys

  hipLaunchKernelGGL(k_plan_header, dim3(1), dim3(64), 0, s, hdr, n, L.ncx, L.ncy);
  if (n > 0) {
    const double W = (p->n_il - 1) * (double)L.d_il, H = (p->n_xl - 1) * (double)L.d_xl;
    const double A = p->aperture_m;
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_plan_local, dim3(blocks), dim3(256), 0, s, src_xy, rcv_xy, n,
                       p->origin_x, p->origin_y, L.ux_x, L.ux_y, L.uy_x, L.uy_y, -A, W + A, -A,
                       H + A, (double)L.cell_x0, (double)L.cell_y0, (double)L.cell_w,
                       (double)L.cell_h, L.ncx, L.ncy, L.n_bands, 2.0 * A, p->min_offset_m,
                       p->max_offset_m, local, offs_local, keys_in, ids_in);
    size_t tmp_bytes = L.sort_tmp_bytes;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(base + L.off_sort_tmp, tmp_bytes, keys_in,
                                                      keys_out, ids_in, ids_out, (int)n, 0,
                                                      L.key_bits, s);
    if (e != hipSuccess) return set_hip_error("radix sort", e);
    hipLaunchKernelGGL(k_plan_gather, dim3(blocks), dim3(256), 0, s, local, offs_local, ids_out, n, geom,
                       offs);
  }
  int cblocks = (int)((n_cells + 1 + 255) / 256);
  hipLaunchKernelGGL(k_plan_cells, dim3(cblocks), dim3(256), 0, s, keys_out, n, n_cells, cells,
                     hdr);
  return check_launch("plan_build");
}

// ---------------------------------------------------------------------------
// FP32 -> BF16 / FP16 with one FP32 scale per trace (scale = max|trace|).

__device__ inline uint16_t f32_to_bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}

__global__ void k_quantize(const float* __restrict__ x, int32_t n_samples, int32_t dtype,
                           void* __restrict__ out, float* __restrict__ scale) {
  const int64_t tr = blockIdx.x;
  const float* row = x + tr * (int64_t)n_samples;
  __shared__ float red[kThreads / 64];
  float m = 0.f;
  for (int i = threadIdx.x; i < n_samples; i += blockDim.x) m = fmaxf(m, fabsf(row[i]));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t = fmaxf(t, red[w]);
    red[0] = t;
  }
  __syncthreads();
  const float smax = red[0];
  const float sc = smax > 0.f ? smax : 1.f;
  const float inv = 1.0f / sc;
  if (threadIdx.x == 0) scale[tr] = sc;
  uint16_t* o = (uint16_t*)out + tr * (int64_t)n_samples;
  for (int i = threadIdx.x; i < n_samples; i += blockDim.x) {
    float v = row[i] * inv;
    if (dtype == PSTM_BF16) {
      o[i] = f32_to_bf16_rne(v);
    } else {
      _Float16 h = (_Float16)v;  // round to nearest even
      o[i] = *(uint16_t*)&h;
    }
  }
}

int quantize(const float* x, int64_t n_traces, int32_t n_samples, int32_t dtype, void* out,
             float* scale, hipStream_t s) {
  if (n_traces <= 0) return PSTM_OK;
  hipLaunchKernelGGL(k_quantize, dim3((unsigned)n_traces), dim3(kThreads), 0, s, x, n_samples,
                     dtype, out, scale);
  return check_launch("quantize_traces");
}

}  // namespace pstm
