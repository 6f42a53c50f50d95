// pstm_abi.hip -- extern "C" entry points of libpstm_hip.so (see
// include/pstm_abi.h): parameter validation, plan layout, error reporting and
// kernel dispatch.  No entry point allocates device memory or synchronises,
// except pstm_plan_kept (which reads one value back).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "pstm_internal.h"

namespace pstm {

static thread_local std::string g_last_error;

int set_error(int status, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return status;
}

int set_hip_error(const char* what, hipError_t e) {
  return set_error(PSTM_ERR_HIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_hip_error(what, e);
  return PSTM_OK;
}

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

static int validate(const pstm_params_t* p) {
  if (!p) return set_error(PSTM_ERR_INVALID, "params is NULL");
  if (p->n_traces < 0 || p->n_traces > 0x7fffffffLL)
    return set_error(PSTM_ERR_INVALID, "n_traces out of range: %lld", (long long)p->n_traces);
  if (p->n_samples < 2) return set_error(PSTM_ERR_INVALID, "n_samples must be >= 2");
  if (p->n_il < 1 || p->n_xl < 1 || p->n_tau < 1)
    return set_error(PSTM_ERR_INVALID, "image dimensions must be >= 1");
  if (!(p->dt > 0) || !(p->dtau > 0)) return set_error(PSTM_ERR_INVALID, "dt and dtau must be > 0");
  if (!(p->tau0 >= 0)) return set_error(PSTM_ERR_INVALID, "tau0 must be >= 0");
  if (!(p->aperture_m > 0)) return set_error(PSTM_ERR_INVALID, "aperture_m must be > 0");
  if (!(p->max_angle_deg > 0 && p->max_angle_deg <= 90))
    return set_error(PSTM_ERR_INVALID, "max_angle_deg must be in (0, 90]");
  if (p->dtype < PSTM_F32 || p->dtype > PSTM_F16) return set_error(PSTM_ERR_INVALID, "bad dtype %d", p->dtype);
  if (p->weight_mode < PSTM_WEIGHT_FULL || p->weight_mode > PSTM_WEIGHT_OBLIQUITY)
    return set_error(PSTM_ERR_INVALID, "bad weight_mode %d", p->weight_mode);
  if (p->vel_mode != PSTM_VEL_1D && p->vel_mode != PSTM_VEL_3D)
    return set_error(PSTM_ERR_INVALID, "bad vel_mode %d", p->vel_mode);
  const int known = PSTM_FLAG_AA | PSTM_FLAG_VTI | PSTM_FLAG_KAHAN | PSTM_FLAG_TAPER |
                    PSTM_FLAG_ZERO_OUTSIDE | PSTM_FLAG_TIME_APERTURE;
  if (p->flags & ~known) return set_error(PSTM_ERR_INVALID, "unknown flag bits 0x%x", p->flags & ~known);
  if ((p->flags & PSTM_FLAG_TAPER) && !(p->taper_width > 0 && p->taper_width < 1))
    return set_error(PSTM_ERR_INVALID, "taper_width must be in (0, 1)");
  if (!(p->min_offset_m >= 0) || !(p->max_offset_m == p->max_offset_m))
    return set_error(PSTM_ERR_INVALID, "min_offset_m must be >= 0 and max_offset_m a number");
  if (p->max_offset_m > 0 && !(p->min_offset_m < p->max_offset_m))
    return set_error(PSTM_ERR_INVALID, "min_offset_m (%g) must be below max_offset_m (%g)", p->min_offset_m,
                     p->max_offset_m);
  const double dil = hypot(p->il_dx, p->il_dy), dxl = hypot(p->xl_dx, p->xl_dy);
  if (!(dil > 0) || !(dxl > 0)) return set_error(PSTM_ERR_INVALID, "grid steps must be non-zero");
  if (fabs(p->il_dx * p->xl_dx + p->il_dy * p->xl_dy) > 1e-6 * dil * dxl)
    return set_error(PSTM_ERR_INVALID, "inline and crossline steps must be orthogonal");
  int pad = 0;
  if (p->flags & PSTM_FLAG_AA) {
    if (!(p->aa_max >= 1 && p->aa_max <= 256)) return set_error(PSTM_ERR_INVALID, "aa_max must be in [1, 256]");
    if (!(p->aa_scale >= 0)) return set_error(PSTM_ERR_INVALID, "aa_scale must be >= 0");
    pad = (int)ceil(p->aa_max) + 2;
  }
  // Any record length: a trace window longer than a kernel form's LDS share
  // is staged in segments.  The bound keeps window and sample indices far
  // from the 32-bit range.
  (void)pad;
  if (p->n_samples > (1 << 24))
    return set_error(PSTM_ERR_UNSUPPORTED, "n_samples %d above 16,777,216", p->n_samples);
  if ((int64_t)p->n_il * p->n_xl * (int64_t)p->n_tau > ((int64_t)1 << 40))
    return set_error(PSTM_ERR_INVALID, "image too large");
  return PSTM_OK;
}

int compute_layout(const pstm_params_t* p, PlanLayout* L) {
  int rc = validate(p);
  if (rc != PSTM_OK) return rc;
  PlanLayout l = {};
  l.n_traces = p->n_traces;
  const double dil = hypot(p->il_dx, p->il_dy), dxl = hypot(p->xl_dx, p->xl_dy);
  l.ux_x = p->il_dx / dil;
  l.ux_y = p->il_dy / dil;
  l.uy_x = p->xl_dx / dxl;
  l.uy_y = p->xl_dy / dxl;
  l.d_il = (float)dil;
  l.d_xl = (float)dxl;
  // Spatial cells of 4 x 4 image columns, independent of the tile shape a
  // kernel uses (a tile looks up every cell its reach overlaps).
  const double cw = 4 * dil, ch = 4 * dxl, A = p->aperture_m;
  const double W = (p->n_il - 1) * dil, H = (p->n_xl - 1) * dxl;
  const double x0 = -ceil(A / cw) * cw, y0 = -ceil(A / ch) * ch;
  const double ncx = floor((W + A - x0) / cw) + 1, ncy = floor((H + A - y0) / ch) + 1;
  if (ncx * ncy * kOffsetBands > (double)(1 << 30))
    return set_error(PSTM_ERR_UNSUPPORTED, "aperture/grid give too many cells (%.0f)", ncx * ncy);
  l.ncx = (int32_t)ncx;
  l.ncy = (int32_t)ncy;
  l.cell_x0 = (float)x0;
  l.cell_y0 = (float)y0;
  l.cell_w = (float)cw;
  l.cell_h = (float)ch;
  l.n_bands = kOffsetBands;
  const uint32_t n_cells = (uint32_t)l.ncx * (uint32_t)l.ncy * (uint32_t)l.n_bands;  // sort keys
  int bits = 1;
  while (bits < 32 && ((uint64_t)1 << bits) <= n_cells) ++bits;
  l.key_bits = bits;
  const size_t n = (size_t)p->n_traces;
  size_t off = 0;
  l.off_header = off;     off = align_up(off + sizeof(PlanHeader));
  l.off_keys_in = off;    off = align_up(off + 4 * n);
  l.off_keys_out = off;   off = align_up(off + 4 * n);
  l.off_ids_in = off;     off = align_up(off + 4 * n);
  l.off_ids_out = off;    off = align_up(off + 4 * n);
  l.off_local = off;      off = align_up(off + 16 * n);
  l.off_geom = off;       off = align_up(off + 16 * n);
  l.off_offs_local = off; off = align_up(off + 4 * n);
  l.off_offs = off;       off = align_up(off + 4 * n);
  l.off_cell_start = off; off = align_up(off + 4 * ((size_t)n_cells + 1));
  l.sort_tmp_bytes = n > 0 ? sort_temp_bytes((int64_t)n, bits) : 0;
  l.off_sort_tmp = off;   off = align_up(off + l.sort_tmp_bytes);
  l.total_bytes = off;
  *L = l;
  return PSTM_OK;
}

// Tile shape of a call: 2D lines use 8 columns along the line; 3D images use
// 4 x 2 columns (8 per lane: ~130 VGPRs, 3 waves per SIMD on the fast path;
// 4 x 4 tiles need ~240 VGPRs and measured slower).
static void choose_shape(const pstm_params_t* p, int* tx, int* ty) {
  if (p->n_xl == 1) { *tx = 8; *ty = 1; return; }
  if (p->n_il == 1) { *tx = 1; *ty = 8; return; }
  *tx = 4;
  *ty = 2;
}

void fill_tile_args(const pstm_params_t* p, const PlanLayout& L, const void* ws, TileArgs* a) {
  TileArgs t = {};
  choose_shape(p, &t.tile_x, &t.tile_y);
  t.n_samples = p->n_samples;
  t.n_il = p->n_il;
  t.n_xl = p->n_xl;
  t.n_tau = p->n_tau;
  t.n_tiles_x = (p->n_il + t.tile_x - 1) / t.tile_x;
  t.n_tiles_y = (p->n_xl + t.tile_y - 1) / t.tile_y;
  t.n_tblocks = (p->n_tau + kTileTau - 1) / kTileTau;
  t.inv_dt = (float)(1.0 / p->dt);
  t.t0_over_dt = (float)(p->t0 / p->dt);
  t.umax = (float)(p->n_samples - 1.001);
  t.u_last = (float)(p->n_samples - 1);
  t.tau0 = (float)p->tau0;
  t.dtau = (float)p->dtau;
  t.d_il = L.d_il;
  t.d_xl = L.d_xl;
  t.aperture = (float)p->aperture_m;
  t.tan_th = (float)tan(p->max_angle_deg * M_PI / 180.0);
  t.taper_width = (float)p->taper_width;
  t.aa_scale_over_dt = (float)(p->aa_scale / p->dt);
  t.aa_max = (float)p->aa_max;
  t.th_max = (float)(p->max_angle_deg * M_PI / 180.0);
  t.tan_taper = (float)tan(p->max_angle_deg * M_PI / 180.0 * (1.0 - p->taper_width));
  t.off_min = (float)p->min_offset_m;
  t.off_max = p->max_offset_m > 0 ? (float)p->max_offset_m : 0.f;
  t.aa_pad = (p->flags & PSTM_FLAG_AA) ? (int)ceil(p->aa_max) + 2 : 0;
  t.flags = p->flags;
  t.weight_mode = p->weight_mode;
  t.vel_mode = p->vel_mode;
  t.ncx = L.ncx;
  t.ncy = L.ncy;
  t.n_bands = L.n_bands;
  t.band0 = 0;
  t.band1 = L.n_bands;
  t.cell_x0 = L.cell_x0;
  t.cell_y0 = L.cell_y0;
  t.cell_w = L.cell_w;
  t.cell_h = L.cell_h;
  const char* base = (const char*)ws;
  t.geom = (const float4*)(base + L.off_geom);
  t.ids = (const int32_t*)(base + L.off_ids_out);
  t.cell_start = (const int32_t*)(base + L.off_cell_start);
  t.offs = (const float*)(base + L.off_offs);
  *a = t;
}

}  // namespace pstm

using namespace pstm;

extern "C" {

int pstm_abi_version(void) { return PSTM_ABI_VERSION; }

const char* pstm_last_error(void) { return g_last_error.c_str(); }

int pstm_validate(const pstm_params_t* p) {
  g_last_error.clear();
  return validate(p);
}

size_t pstm_plan_bytes(const pstm_params_t* p) {
  g_last_error.clear();
  PlanLayout L;
  if (compute_layout(p, &L) != PSTM_OK) return 0;
  return L.total_bytes;
}

int pstm_plan_build(const pstm_params_t* p, const double* src_xy, const double* rcv_xy, void* ws,
                    size_t ws_bytes, void* stream) {
  g_last_error.clear();
  PlanLayout L;
  int rc = compute_layout(p, &L);
  if (rc != PSTM_OK) return rc;
  if (!ws || ws_bytes < L.total_bytes)
    return set_error(PSTM_ERR_WORKSPACE, "workspace needs %zu bytes, got %zu", L.total_bytes, ws_bytes);
  if (p->n_traces > 0 && (!src_xy || !rcv_xy)) return set_error(PSTM_ERR_INVALID, "NULL coordinates");
  return plan_build(p, L, src_xy, rcv_xy, ws, (hipStream_t)stream);
}

int64_t pstm_plan_kept(const void* ws, void* stream) {
  g_last_error.clear();
  if (!ws) return set_error(PSTM_ERR_WORKSPACE, "NULL workspace");
  PlanHeader h;
  hipError_t e = hipMemcpyAsync(&h, ws, sizeof(h), hipMemcpyDeviceToHost, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) return set_hip_error("plan_kept", e);
  if (h.magic != kPlanMagic) return set_error(PSTM_ERR_WORKSPACE, "workspace holds no plan");
  return h.n_kept;
}

int pstm_quantize_traces(const float* traces, int64_t n_traces, int32_t n_samples, int32_t dtype,
                         void* out, float* scale, void* stream) {
  g_last_error.clear();
  if (dtype != PSTM_BF16 && dtype != PSTM_F16)
    return set_error(PSTM_ERR_INVALID, "quantize target must be BF16 or F16");
  if (n_traces < 0 || n_samples < 1) return set_error(PSTM_ERR_INVALID, "bad sizes");
  if (n_traces > 0 && (!traces || !out || !scale)) return set_error(PSTM_ERR_INVALID, "NULL pointer");
  return quantize(traces, n_traces, n_samples, dtype, out, scale, (hipStream_t)stream);
}

static int prep(const pstm_params_t* p, const void* ws, PlanLayout* L, TileArgs* a) {
  g_last_error.clear();
  int rc = compute_layout(p, L);
  if (rc != PSTM_OK) return rc;
  if (!ws) return set_error(PSTM_ERR_WORKSPACE, "NULL workspace");
  fill_tile_args(p, *L, ws, a);
  return PSTM_OK;
}

int pstm_plan_bands(const pstm_params_t* p) {
  g_last_error.clear();
  PlanLayout L;
  const int rc = compute_layout(p, &L);
  return rc != PSTM_OK ? rc : L.n_bands;
}

int pstm_forward(const pstm_params_t* p, const void* ws, const void* traces,
                 const float* trace_scale, const float* vel, const float* eta, float* image,
                 float* image_comp, float* fold, void* stream) {
  const int nb = pstm_plan_bands(p);
  if (nb < 0) return nb;
  return pstm_forward_bands(p, ws, 0, nb, traces, trace_scale, vel, eta, image, image_comp, fold, stream);
}

// Forward over offset bands [band_begin, band_end): from the traces, or with
// int_a/int_c from their integrals (antialias).
static int forward_common(const pstm_params_t* p, const void* ws, int32_t band_begin, int32_t band_end,
                          const void* traces, const float* trace_scale, const double* int_a, const float* int_c,
                          const float* vel, const float* eta, float* image, float* image_comp, float* fold,
                          void* stream) {
  PlanLayout L;
  TileArgs a;
  int rc = prep(p, ws, &L, &a);
  if (rc != PSTM_OK) return rc;
  if (band_begin < 0 || band_end < band_begin || band_end > L.n_bands)
    return set_error(PSTM_ERR_INVALID, "offset bands [%d, %d) outside [0, %d)", band_begin, band_end, L.n_bands);
  a.band0 = band_begin;
  a.band1 = band_end;
  if (!vel || !image || (p->n_traces > 0 && !traces && !int_a))
    return set_error(PSTM_ERR_INVALID, "NULL traces, vel or image");
  if ((p->flags & PSTM_FLAG_VTI) && !eta) return set_error(PSTM_ERR_INVALID, "PSTM_FLAG_VTI needs eta");
  a.traces = traces;
  a.scale = trace_scale;
  a.int_a = int_a;
  a.int_c = int_c;
  a.vel = vel;
  a.eta = eta;
  a.image = image;
  a.comp = image_comp;
  a.fold = fold;
  if (p->n_traces == 0) return PSTM_OK;
  return launch_forward(p, L, a, (hipStream_t)stream);
}

int pstm_forward_bands(const pstm_params_t* p, const void* ws, int32_t band_begin, int32_t band_end,
                       const void* traces, const float* trace_scale, const float* vel, const float* eta,
                       float* image, float* image_comp, float* fold, void* stream) {
  return forward_common(p, ws, band_begin, band_end, traces, trace_scale, nullptr, nullptr, vel, eta, image,
                        image_comp, fold, stream);
}

int pstm_integrate_traces(const pstm_params_t* p, const void* traces, const float* trace_scale, double* int_a,
                          float* int_c, void* stream) {
  int rc = validate(p);
  if (rc != PSTM_OK) return rc;
  if (p->n_traces > 0 && (!traces || !int_a || !int_c))
    return set_error(PSTM_ERR_INVALID, "NULL traces or integral buffers");
  if (p->n_traces == 0) return PSTM_OK;
  return launch_integrate(p, traces, trace_scale, int_a, int_c, (hipStream_t)stream);
}

int pstm_forward_integrated(const pstm_params_t* p, const void* ws, int32_t band_begin, int32_t band_end,
                            const double* int_a, const float* int_c, const float* vel, const float* eta,
                            float* image, float* image_comp, float* fold, void* stream) {
  if (p && !(p->flags & PSTM_FLAG_AA))
    return set_error(PSTM_ERR_INVALID, "pstm_forward_integrated is the antialias forward (PSTM_FLAG_AA)");
  if (p && p->n_traces > 0 && (!int_a || !int_c)) return set_error(PSTM_ERR_INVALID, "NULL integral buffers");
  return forward_common(p, ws, band_begin, band_end, nullptr, nullptr, int_a, int_c, vel, eta, image, image_comp,
                        fold, stream);
}

int pstm_adjoint(const pstm_params_t* p, const void* ws, const float* image, const float* vel,
                 const float* eta, float* traces_out, void* stream) {
  PlanLayout L;
  TileArgs a;
  int rc = prep(p, ws, &L, &a);
  if (rc != PSTM_OK) return rc;
  if (!vel || !image || (p->n_traces > 0 && !traces_out))
    return set_error(PSTM_ERR_INVALID, "NULL image, vel or traces_out");
  if ((p->flags & PSTM_FLAG_VTI) && !eta) return set_error(PSTM_ERR_INVALID, "PSTM_FLAG_VTI needs eta");
  a.vel = vel;
  a.eta = eta;
  a.image_in = image;
  a.traces_out = traces_out;
  if (p->n_traces == 0) return PSTM_OK;
  return launch_adjoint(p, L, a, (hipStream_t)stream);
}

int pstm_vel_grad(const pstm_params_t* p, const void* ws, const void* traces,
                  const float* trace_scale, const float* vel, const float* image_grad,
                  float* vel_grad, void* stream) {
  PlanLayout L;
  TileArgs a;
  int rc = prep(p, ws, &L, &a);
  if (rc != PSTM_OK) return rc;
  if (p->flags & (PSTM_FLAG_AA | PSTM_FLAG_VTI | PSTM_FLAG_ZERO_OUTSIDE))
    return set_error(PSTM_ERR_UNSUPPORTED, "vel_grad supports isotropic, non-antialiased, clamped migration only");
  if (p->vel_mode != PSTM_VEL_3D) return set_error(PSTM_ERR_UNSUPPORTED, "vel_grad needs PSTM_VEL_3D");
  if (!vel || !image_grad || !vel_grad || (p->n_traces > 0 && !traces))
    return set_error(PSTM_ERR_INVALID, "NULL pointer argument");
  a.traces = traces;
  a.scale = trace_scale;
  a.vel = vel;
  a.image_in = image_grad;
  a.vel_grad = vel_grad;
  if (p->n_traces == 0) return PSTM_OK;
  return launch_vel_grad(p, L, a, (hipStream_t)stream);
}

}  // extern "C"
