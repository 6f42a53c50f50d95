// pstm_common.h -- shared host/device definitions of the gfx950 PSTM library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pstm_abi.h"

namespace pstm {

constexpr uint64_t kPlanMagic = 0x50535450'4C414E31ull;  // "PSTPLAN1"

// Forward/adjoint tile: TILE_TAU output times per workgroup (one per thread),
// a TX x TY block of image columns held in registers by every thread.
constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kTileTau = kThreads;

// Margin (samples) added on each side of a staged window beyond the
// rigorous traveltime bounds (absorbs FP32 rounding of the bounds).
constexpr int kWindowMargin = 2;

// Traces per sub-batch: at high occupancy the small LDS share holds the
// windows of only ~10 traces, so the per-batch arrays shrink with it.
constexpr int batch_of(int occ) { return occ >= 6 ? 16 : occ >= 4 ? 32 : 64; }
// The plan orders traces by (offset band, midpoint cell) and the tile
// operators run one launch per band, so the grid reads traces of one band at
// a time (a working set that stays in the caches; measured 5-7 % faster than
// one launch over all offsets on the 1M-trace, 25-class workload).
constexpr int kOffsetBands = 16;
// Adjoint fixed-point taps: the FP32 form rounds every tap at 2^-30 of the
// workgroup's largest per-lane bound G.  It is taken when the smallest nonzero
// per-lane bound is at least G / kNarrowSpan, so that a lane's taps are rounded
// at 2^-22 of its own bound or finer (the FP32 output traces round at 2^-24 of
// each sample); wider spans (the shallowest time block under the FULL weight,
// or image tiles with near-zero columns) take FP64 taps rounded at 2^-47 G.
constexpr float kNarrowSpan = 256.f;
// Traces summed in FP32 per wave before the partial sums join the running
// (compensated) sums: a flush costs about one trace's arithmetic per column,
// so flushing after every small sub-batch of the high-occupancy forms cost 4 %
// of the forward time.
constexpr int kFlushTraces = 32;
// Candidate queue entries: a batch's remainder (< MB) plus one chunk of kThreads.
constexpr int queue_cap(int mb) { return ((mb - 1 + kThreads) + 15) & ~15; }

// LDS budget of one tile workgroup for OCC (2 to 8) workgroups per CU
// (160 KB per CU, allocated in 512-byte granules).  Three workgroups of
// 54272 bytes (159 KB) were measured to fit only two per CU; 53760 bytes
// (157.5 KB for three) fit three; the others keep the same 1.5 KB of slack.
constexpr int lds_budget(int occ) {
  return occ >= 8 ? 19968 : occ >= 7 ? 22528 : occ >= 6 ? 26112 : occ >= 5 ? 31744 : occ >= 4 ? 39936 : occ >= 3 ? 53760 : 80896;
}
static_assert(lds_budget(3) % 512 == 0 && lds_budget(2) % 512 == 0, "512-byte LDS granules");

// Tile-kernel LDS besides the staged sample windows (MB traces per batch).
template <int MB>
struct TileShared {
  int32_t q_j[queue_cap(MB)];   // candidate queue: sorted trace index
  int32_t q_w0[queue_cap(MB)];  // first staged sample
  int32_t q_len[queue_cap(MB)]; // staged window length
  uint8_t q_nc[queue_cap(MB)];  // bit w: no time of wave w reads past either end of the record
  int32_t b_base[MB];           // window-element index of sample 0 of the trace (off - w0)
  int32_t b_off[MB];            // window-element offset of the window
  int32_t scan[kWaves];
  int32_t queue_n;
  int32_t nb;
  float hwave[kWaves];          // max hlim^2 of each wave (over its columns and times)
  float qmin[kWaves], qmax[kWaves], tqa[kWaves], tqb[kWaves];
  float emaxw[kWaves];          // per wave: max(eta, 0) over its columns and times (VTI)
  float eminw[kWaves];          // per wave: min(eta, 0)
  float gmax[kWaves];           // adjoint: per-wave bound of one contribution
  float gmin[kWaves];           // adjoint: per-wave minimum of the per-time bounds
};

// Per sub-batch (trace, column) quantities of a tile of NC columns.
template <int NC, bool AA, int MB>
struct PairShared {
  float2 h2[MB][NC];            // squared source / receiver distances (m^2)
  float hm[MB][NC];             // max of the two
  int32_t act[MB];              // bit w: the trace reaches some time of wave w; bit 4 + w: it stays
                                // inside the record in wave w and (forward) passes every time and
                                // column of it; bit 8 + w (forward): passes them all but clamps;
                                // (plain adjoint) bit 8 + w: inside the record and passes them all,
                                // bit 12 + w: passes them all but clamps
  float hcol[kWaves][NC];       // per wave and column: max hlim^2 over the wave's times
  float hcolmin[kWaves][NC];    // per wave and column: min hlim^2 over the wave's times
  float4 g[AA ? MB : 1][AA ? NC : 1];  // antialias: scaled (s - x, r - x) components
};

// Staged-window capacity in floats and in window elements.  A window
// element is one FP32 sample, or with the antialias filter the FP64 double
// integral plus the FP32 single integral (three floats).
template <int NC, bool AA, int OCC>
constexpr int sample_cap_floats() {
  return ((lds_budget(OCC) - (int)sizeof(TileShared<batch_of(OCC)>) -
           (int)sizeof(PairShared<NC, AA, batch_of(OCC)>)) / 4) & ~63;
}
template <int NC, bool AA, int OCC>
constexpr int sample_cap_elems() {
  return AA ? sample_cap_floats<NC, AA, OCC>() / 3 : sample_cap_floats<NC, AA, OCC>();
}
// Adjoint windows in 64-bit fixed point (FX) take two floats per element
// without antialias (with antialias the element already holds 8 bytes).
template <int NC, bool AA, int OCC, bool FX>
constexpr int window_cap_elems() {
  return (FX && !AA) ? sample_cap_floats<NC, AA, OCC>() / 2 : sample_cap_elems<NC, AA, OCC>();
}
// The antialias forward at 3 or more workgroups per CU takes 16 bytes per
// window element (FP64 double integral, and FP32 single integral in an
// 8-byte slot of a second array).
template <int NC, int OCC = 3>
constexpr int aa_fast_cap_elems() {
  return sample_cap_floats<NC, true, OCC>() / 4;
}
// The plain forward at 4 or more workgroups per CU stores 8-byte window
// elements: the sample and its forward difference.
template <int OCC>
constexpr int pair_cap_elems() {
  return sample_cap_floats<8, false, OCC>() / 2;
}
// Word 3 of a raw buffer resource (gfx9: 32-bit data format, no swizzle,
// stride 0): loads at byte offsets at or past num_records return 0.
constexpr int kRawBufferWord3 = 0x00020000;
// Workgroups per CU of the kernel forms that take any admissible trace
// length: 3 for the plain tiles and the velocity gradient, 2 for antialias.
// Traces whose whole window fits a smaller LDS share run at higher
// occupancy: the plain forward and the plain adjoint at 6, then 4 (7 first
// with a velocity per output time), the antialias forward at 4 (else 3).
constexpr int default_occ(bool aa) { return aa ? 2 : 3; }
// Capacity every call can rely on (8-column tiles).
constexpr int min_cap_elems(bool aa) {
  return aa ? sample_cap_elems<8, true, default_occ(true)>() : sample_cap_elems<8, false, default_occ(false)>();
}
// Elements of a bank-skewed adjoint window of `len` samples (one spare
// element before the window and one before every 32nd sample).
__host__ __device__ constexpr int skew_len(int len) { return len + (len >> 5) + 2; }

// Window length (elements) of a whole trace with its padding.
inline int64_t whole_window(const pstm_params_t* p, int aa_pad) {
  return (int64_t)p->n_samples + 2 * (aa_pad + kWindowMargin) + 2;
}


// Plan layout: everything the kernels need about the trace binning, derived
// deterministically from the job parameters on the host.
struct PlanLayout {
  int64_t n_traces;
  int32_t ncx, ncy;      // spatial cells of 4 x 4 image columns (along il, along xl)
  int32_t n_bands;       // offset bands of width 2 aperture / n_bands (key = band * cells + cell)
  int32_t key_bits;      // radix-sort key width
  float cell_x0, cell_y0;  // local coordinate of the cell grid corner (m)
  float cell_w, cell_h;    // cell size (m)
  float d_il, d_xl;        // column spacing (m)
  double ux_x, ux_y;       // unit vector of the inline axis (world)
  double uy_x, uy_y;       // unit vector of the crossline axis (world)
  // byte offsets inside the workspace
  size_t off_header, off_keys_in, off_keys_out, off_ids_in, off_ids_out;
  size_t off_local, off_geom, off_cell_start, off_sort_tmp;
  size_t off_offs_local, off_offs;  // per-trace offsets (FP32 of the FP64 distance), unsorted / sorted
  size_t sort_tmp_bytes;
  size_t total_bytes;
};

// Device-resident header at workspace offset 0.
struct PlanHeader {
  uint64_t magic;
  int64_t n_traces;
  int64_t n_kept;
  int32_t ncx, ncy;
};

// Arguments of the tile kernels (forward / adjoint / velocity gradient).
struct TileArgs {
  // sizes
  int32_t n_samples, n_il, n_xl, n_tau;
  int32_t tile_x, tile_y;  // image columns per tile along il / xl
  int32_t n_tiles_x, n_tiles_y, n_tblocks;
  // time axes
  float inv_dt, t0_over_dt, umax, u_last;  // umax = n-1.001, u_last = n-1
  float tau0, dtau;
  // geometry
  float d_il, d_xl;
  float aperture, tan_th;
  float taper_width;
  float aa_scale_over_dt, aa_max;
  float th_max;             // max_angle in radians (taper)
  float tan_taper;          // tan(th_max (1 - taper_width)): the angle taper's flat zone ends there
  float off_min, off_max;   // offset window (taper; off_max <= 0: none)
  int32_t aa_pad;           // staged zero padding on each side (samples)
  int32_t flags, weight_mode, vel_mode;
  // plan
  int32_t ncx, ncy, n_bands;
  int32_t band0, band1;      // offset bands [band0, band1) this launch processes
  float cell_x0, cell_y0, cell_w, cell_h;
  const float4* geom;        // sorted local (sa, sb, ra, rb)
  const int32_t* ids;        // sorted -> original trace index
  const int32_t* cell_start; // [n_bands*ncx*ncy + 1]
  const float* offs;         // sorted source-receiver offsets (m)
  // data
  const void* traces;
  const float* scale;
  const float* vel;
  const float* eta;
  float* image;
  float* comp;
  float* fold;
  // adjoint / gradient
  const float* image_in;     // adjoint input image or gradient upstream
  float* traces_out;         // adjoint output
  float* vel_grad;           // gradient output
  // antialias forward with traces integrated once per call (or NULL: each
  // staged window is integrated in LDS): [trace][sample] FP64 double integral
  // A(s) = sum_{k >= s} C(k) and FP32 single integral C(s) = sum_{j <= s} x(j),
  // storage scale applied
  const double* int_a;
  const float* int_c;
};

}  // namespace pstm
